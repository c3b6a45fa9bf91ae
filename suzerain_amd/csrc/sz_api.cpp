// Host-side BSMBSM geometry, the C-ABI counterpart of suzerain_bsmbsm_construct,
// suzerain_bsmbsm_q and suzerain_bsmbsm_qinv (suzerain/bsmbsm.h:104-186).
#include "../../include/suzerain_amd.h"

extern "C" {

sz_bsmbsm sz_bsmbsm_construct(int S, int n, int kl, int ku)
{
    sz_bsmbsm t;
    t.S = S;
    t.n = n;
    t.kl = kl;
    t.ku = ku;
    t.ld = kl + 1 + ku;
    t.N = S * n;
    t.KL = S * (kl + 1) - 1;
    t.KU = S * (ku + 1) - 1;
    t.LD = t.KL + 1 + t.KU;
    return t;
}

int sz_bsmbsm_q(int S, int n, int i) { return (i % S) * n + i / S; }

int sz_bsmbsm_qinv(int S, int n, int i) { return (i % n) * S + i / n; }

}  // extern "C"

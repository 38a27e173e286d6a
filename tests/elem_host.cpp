// Host build of the element arithmetic (fluidity_amd/csrc/element_kernels.h)
// for the CPU unit tests: the same source the gfx950 kernels run, compiled
// with g++, evaluated one element at a time.
#include <cstring>

#include "../fluidity_amd/csrc/element_kernels.h"

using namespace flgpu;

template <int DIM, int LOC, typename R>
static void run(const double* tab_n, const double* tab_dn, const double* w, const double* xdn,
                const double* what, const double* mhat, const double* that, int ngi,
                const double* X /* dim x (dim+1) col-major: X[a + dim*n] */, const int* nodes,
                const double* T, const AsmFlags& fl, const AsmFields& fd, double* mat /* loc x loc col-major */,
                double* rhs) {
    LTab<R> tb;
    std::memset(&tb, 0, sizeof tb);
    for (int k = 0; k < LOC * ngi; k++) tb.n[k] = (R)tab_n[k];
    for (int k = 0; k < LOC * ngi * DIM; k++) tb.dn[k] = (R)tab_dn[k];
    for (int g = 0; g < ngi; g++) tb.w[g] = (R)w[g];
    for (int k = 0; k < (DIM + 1) * DIM; k++) tb.xdn[k] = (R)xdn[k];
    for (int k = 0; k < LOC; k++) tb.what[k] = (R)what[k];
    for (int k = 0; k < LOC * LOC; k++) tb.mhat[k] = (R)mhat[k];
    for (int k = 0; k < 64; k++) tb.that[k] = (R)that[k];
    tb.loc = LOC;
    tb.ngi = ngi;
    tb.lx = DIM + 1;
    Elem<DIM, LOC, R> e;
    for (int k = 0; k < LOC; k++) {
        e.nodes[k] = nodes[k];
        e.T[k] = (R)T[nodes[k]];
    }
    for (int nn = 0; nn < DIM + 1; nn++)
        for (int a = 0; a < DIM; a++) e.X[a][nn] = (R)X[a + DIM * nn];
    for (int i = 0; i < LOC; i++) {
        R vals[LOC], r;
        if (fl.stab)
            row_stab<DIM, LOC, R>(e, tb, fl, fd, i, vals, r);
        else
            row_contrib<DIM, LOC, R>(e, tb, fl, fd, i, vals, r);
        for (int j = 0; j < LOC; j++) mat[i + LOC * j] = (double)vals[j];
        rhs[i] = (double)r;
    }
}

extern "C" int elem_host(int dim, int loc, int fp32, const double* n, const double* dn, const double* w,
                         const double* xdn, const double* what, const double* mhat, const double* that, int ngi,
                         const double* X, const int* nodes, const double* T, const double* U, int u_c,
                         const double* K, int k_c, const double* S, int s_c, const double* B, int b_c,
                         const double* fparams /* dt_theta, beta, nu_scale */,
                         const int* iparams /* 10 flags, stabilisation, nu_bar scheme */,
                         double* mat, double* rhs) {
    AsmFlags fl;
    fl.dt_theta = fparams[0];
    fl.beta = fparams[1];
    fl.have_mass = iparams[0];
    fl.lump = iparams[1];
    fl.adv = iparams[2];
    fl.ibp = iparams[3];
    fl.diff = iparams[4];
    fl.iso = iparams[5];
    fl.src = iparams[6];
    fl.absorp = iparams[7];
    fl.need_div = iparams[8];
    fl.dtt_on = iparams[9];
    fl.stab = iparams[10];
    fl.nu_scheme = iparams[11];
    fl.nu_scale = fparams[2];
    AsmFields fd{T, U, u_c, K, k_c, S, s_c, B, b_c};
#define CASE(D, L)                                                                                     \
    if (dim == D && loc == L) {                                                                        \
        if (fp32)                                                                                      \
            run<D, L, float>(n, dn, w, xdn, what, mhat, that, ngi, X, nodes, T, fl, fd, mat, rhs);     \
        else                                                                                           \
            run<D, L, double>(n, dn, w, xdn, what, mhat, that, ngi, X, nodes, T, fl, fd, mat, rhs);    \
        return 0;                                                                                      \
    }
    CASE(2, 3)
    CASE(2, 6)
    CASE(3, 4)
    CASE(3, 10)
    return -1;
}

// Test probe (built by tests/test_leaves_cpu.py with the host compiler): the
// library's per-face and per-cell physics leaves -- the same source the GPU
// kernels compile (step_cells.h policies over phys_euler.h, phys_mhd.h,
// phys_grmhd_fv.h, phys_greuler.h) -- evaluated on records read from a file, in the layout
// of oracle/ref/ref_leaf.c, so the two can be compared record by record.
//
// usage: leaf_probe PARFILE IN.bin OUT.bin
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "host_setup.h"
#include "step_cells.h"

using namespace disco;

template <class Ph>
static void record(const PhysParams& P, int hydro, const double* in, double* out) {
    constexpr int NQ = Ph::NQ;
    const double* pL = in;
    const double* pR = in + NQ;
    const double* x = in + 2 * NQ;
    const double* n = x + 3;
    double w = n[3], dAdt = n[4];
    int dim = (int)n[5];
    double dU[NQ];
    // the flux kernel's instantiation for runs with HLLC_DAMPCENTER
    FaceEMF e = P.riemann == DISCO_HLLC_DAMPCENTER ? Ph::template riemann<true>(P, pL, pR, x, n, w, dAdt, dim, dU)
                                                   : Ph::riemann(P, pL, pR, x, n, w, dAdt, dim, dU);
    double* o = out;
    for (int q = 0; q < NQ; ++q) *o++ = dU[q];
    *o++ = e.E1;
    *o++ = e.B1;
    *o++ = e.E2;
    *o++ = e.B2;
    double cons[NQ], prim[NQ];
    if (hydro == DISCO_HYDRO_GRMHD) {
        grmhd::Metric g;
        grmhd::kerr_metric(x, P.gr_M, P.gr_a, &g);
        grmhd::prim2cons(P.gamma_law, g, pL, cons, x, 1.37);
    } else if (hydro == DISCO_HYDRO_GRMHD2) {
        grmhd::Metric g;
        grmhd::kerr_metric(x, P.gr_M, P.gr_a, &g);
        grmhd::prim2cons<true>(P.gamma_law, g, pL, cons, x, 1.37);
    } else if (hydro == DISCO_HYDRO_GREULER) {
        grmhd::Metric g;
        grmhd::kerr_metric(x, P.gr_M, P.gr_a, &g);
        greuler::prim2cons(P.gamma_law, g, pL, cons, 1.37);
    } else if (hydro == DISCO_HYDRO_MHD) {
        mhd::prim2cons<NQ>(P, pL, cons, x, 1.37);
    } else {
        euler::prim2cons<NQ>(P, pL, cons, x, 1.37);
    }
    for (int q = 0; q < NQ; ++q) prim[q] = pL[q] * 1.001;
    Ph::cons2prim(P, cons, prim, x, 1.37);
    for (int q = 0; q < NQ; ++q) *o++ = cons[q];
    for (int q = 0; q < NQ; ++q) *o++ = prim[q];
    double xp[3] = {x[0] + 0.05, x[1] + 0.02, x[2] + 0.03};
    double xm[3] = {x[0] - 0.05, x[1] - 0.02, x[2] - 0.03};
    *o++ = Ph::mindt(P, pL, w, xp, xm);
    double src[NQ];
    for (int q = 0; q < NQ; ++q) src[q] = 0.0;
    Ph::source(P, pL, src, xp, xm, 0.7);
    for (int q = 0; q < NQ; ++q) *o++ = src[q];
}

int main(int argc, char** argv) {
    if (argc < 4) return 2;
    std::ifstream pf(argv[1]);
    std::stringstream ss;
    ss << pf.rdbuf();
    disco_params par;
    std::string err;
    if (read_par_text(ss.str(), &par, &err)) {
        std::fprintf(stderr, "%s\n", err.c_str());
        return 1;
    }
    PhysParams P;
    disco_ic ic;
    derive_params(par, &P, &ic);
    FILE* fi = std::fopen(argv[2], "rb");
    FILE* fo = std::fopen(argv[3], "wb");
    if (!fi || !fo) return 1;
    int n = 0;
    if (std::fread(&n, sizeof n, 1, fi) != 1) return 1;
    const int NQ = par.num_c + par.num_n;
    std::vector<double> in(2 * NQ + 9), out(4 * NQ + 5);
    for (int k = 0; k < n; ++k) {
        if (std::fread(in.data(), sizeof(double), in.size(), fi) != in.size()) return 1;
        if (par.hydro == DISCO_HYDRO_GRMHD)
            record<PhysGRMHD>(P, par.hydro, in.data(), out.data());
        else if (par.hydro == DISCO_HYDRO_GREULER)
            record<PhysGREuler>(P, par.hydro, in.data(), out.data());
        else if (par.hydro == DISCO_HYDRO_GRMHD2)
            record<PhysGRMHD2>(P, par.hydro, in.data(), out.data());
        else if (par.hydro == DISCO_HYDRO_MHD && par.num_n == 1)
            record<PhysMHD1>(P, par.hydro, in.data(), out.data());
        else if (par.hydro == DISCO_HYDRO_MHD)
            record<PhysMHD>(P, par.hydro, in.data(), out.data());
        else if (par.num_n == 1)
            record<PhysEuler1>(P, par.hydro, in.data(), out.data());
        else
            record<PhysEuler>(P, par.hydro, in.data(), out.data());
        std::fwrite(out.data(), sizeof(double), out.size(), fo);
    }
    std::fclose(fi);
    std::fclose(fo);
    return 0;
}

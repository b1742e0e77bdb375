// fe_setup.h -- omega-independent host set-up of the DSM engines: grid,
// element integrals, receiver stencils.  Compiled with -ffp-contract=off
// and written in the reference's operation order: the reference's element
// integrals integrate raw monomials in r over thin layers and keep only ~8
// significant digits, so any change of rounding here (FMA contraction,
// reassociation) moves the catalog by ~1e-7.  The l-dependent hot path on
// the GPU does not have this sensitivity.
#pragma once

#include <complex>
#include <string>
#include <vector>

#include "dsm_sgt.h"

namespace dsm {

using cd = std::complex<double>;

// SH set-up: SGTsh-MPI.f90:66-348.
struct ShSetup {
    int nzone = 0;                         // zones above the fluid core
    std::vector<double> vrmin, vrmax, rrho, vsv, vsh, qmu;  // rrho(4,nzone) ...
    double rmin = 0, rmax = 0, omegai = 0, tlen = 0, ratl = 0;
    int maxlmax = 0, imax = 0;
    int r_n = 0, theta_n = 0;
    std::vector<double> r_, theta_deg;     // receivers (km), distances (deg)
    std::vector<double> rrsta;             // rrsta(3,r_n)
    std::vector<int> iista;                // iista(3,r_n), 1-based node indices
    std::vector<double> w0, w1;            // interpolation weights (value, d/dr), (3,r_n)
    double r0 = 0;                         // source radius after calspo
    int intir0 = 0;                        // int(r0*1e3) before calspo
    int nnlayer = 0;
    std::vector<int> nlayer, isp, jsp;
    std::vector<double> ra;
    double spo = 0;
    int spn = 0;
    double gra[3]{}, mu0 = 0;
    std::vector<double> t, h1, h2, h3, h4;  // 4*nnlayer element integrals
    double gt[8]{}, gh1[8]{}, gh2[8]{}, gh3[8]{}, gh4[8]{};
    int nn = 0, ns = 0, ins = 0;
};

// Returns 0 or a DSM_SGT_E* code with *err set.
int sh_setup(const dsm_sgt_input_t &in, ShSetup &s, std::string &err);

// Per-frequency band pieces a0, a2 (band storage a(2,nn): a(1,j) = A(j-1,j),
// a(2,j) = A(j,j)) and the scalars the l loop needs: SGTsh-MPI.f90:400-415.
struct ShOmega {
    double omega = 0;
    int lsuf = 0;
    cd comega2, coef_spn;
    std::vector<cd> a0, a2;   // 2*nn each
};
void sh_omega_setup(const ShSetup &s, int i, ShOmega &o);

// Interpolation weights of the reference's 3-node Taylor stencil
// (others.f90 interpolate + glu): weights b with u = sum b_i g_i.
void stencil_weights(double rsta, const double rr[3], int nderiv, double w[3]);

}  // namespace dsm

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

// ---- P-SV set-up: SGTpsv-MPI.f90:120-480 ---------------------------------
// Element integrals of one zone (calmat.f90 calmatc/caltl/calhl/calh5/
// calhm1/calhm2 and their transposes), stored per zone.
struct PsvZoneFE {
    bool solid = true;
    int nlayer = 0;
    // solid zones, 4*nlayer each
    std::vector<double> t, h1x, h1y, h1z, h2L, h2N, h3ax, h3ay, h3az, h4aL, h4aN, h5ax, h5ay, h5az, h6aL, h6aN, h3x,
        h3y, h3z, h4L, h4N, h5x, h5y, h5z, h6L, h6N, h7x, h7y, h7z, h8L, h8N;
    // solid zones, 4*(nlayer+1) each: h?m?(-2:1 or -1:2, nlayer+1)
    std::vector<double> h3mx, h3my, h3mz, h5mx, h5my, h5mz, h4m1L, h4m1N, h4m2L, h4m2N, h6m1L, h6m1N, h6m2L, h6m2N;
    // fluid zones, 4*nlayer each
    std::vector<double> p1, p2, p3;
};

struct PsvSetup {
    int nzone = 0, ndc = 0;
    std::vector<int> iphase, nlayer;  // 1 solid, 2 fluid
    std::vector<double> vrmin, vrmax, rrho, vpv, vph, vsv, vsh, eta, qmu, qkappa;
    double rmin = 0, rmax = 0, omegai = 0, tlen = 0, ratl = 0;
    int maxlmax = 0, imax = 0;
    int r_n = 0, theta_n = 0;
    std::vector<double> r_, theta_deg, rrsta, w0, w1, lambda, qkp;
    std::vector<int> iista, istazone, ksta, jsta, fluid_rec;
    double r0 = 0, spo = 0;
    int intir0 = 0, spn = 0;
    int nnlayer = 0, nslay = 0, nllay = 0;
    std::vector<double> ra;
    std::vector<int> isp, kkdr;
    std::vector<PsvZoneFE> fe;
    double ecC0 = 0, ecF0 = 0, ecL0 = 0;
    int nn = 0;       // band rows (SGTpsv-MPI.f90:634)
    int row0 = 1;     // first solved row: 3 when the model reaches r = 0
    int jtmp = 0;     // first source row (1-based, SGTpsv-MPI.f90:643-648)
    int mtmp = 0;     // source layer: ra(mtmp), ra(mtmp+1)
    // band assembly (cala + calbc, calmat.f90:747-809): for every column j
    // and band row ii (a(ii,j) = A(j-4+ii, j)) the kind of the entry --
    // 0 zero, 1 a0 + l(l+1) a2, 2 sqrt(l(l+1)) a1, 3 calbc constant -- and
    // whether a second zone adds to it (the node shared by two solid zones)
    std::vector<int> term_kind;       // [nn][4]
    std::vector<int> term_two;        // [nn][4]
    // l = 0 branch (rea2, calmat.f90:1247-1315): node vector of nn0 entries
    int nn0 = 0, row0_0 = 1;
    std::vector<int> node_row;        // [nn0] band column of the node's U (or P) row, 1-based
    std::vector<int> node_c1;         // [nn0] band row ii of c(1,i) (0: none)
    int js0 = 0;                      // node index (0-based) of the source row jtmp
    std::vector<int> jsta0;           // [r_n] first d index of the receiver stencil (1-based)
};

// Returns 0 or a DSM_SGT_E* code with *err set.
int psv_setup(const dsm_sgt_input_t &in, PsvSetup &s, std::string &err);

// The input's distances in degrees, as both set-ups take them: the list, or
// thetamin + (i - 1) thetadelta for i = 1 .. (thetamax - thetamin) / thetadelta + 1
// (SGTpsv-MPI.f90 / SGTsh-MPI.f90 theta grid); returns their count.
int theta_grid(const dsm_sgt_input_t &in, std::vector<double> &deg);

// Per-frequency data (SGTpsv-MPI.f90:566-625): the band pieces of every
// column, the excitation's RK tables and the source-zone moduli.
struct PsvOmega {
    double omega = 0;
    int lsuf = 0;
    cd coef1_spn, coef2_spn, womega;
    std::vector<cd> anum, bnum;       // 160 each (anum(4,4,10))
    // per column j (0-based) and band row ii: up to two contributions
    // (P0,P1), (Q0,Q1); kinds in PsvSetup::term_kind
    std::vector<cd> P0, P1, Q0, Q1;   // [nn][4]
    std::vector<cd> fluidK;           // [r_n]: (omega - i omegai)/3 / lambda~ (calupfluid)
    // scalars of the band assembly: per zone the moduli factors coef1,
    // coef2 and 1/coef2 (calcoef), and (omega - i omegai)^2
    std::vector<cd> zc1, zc2, zcf;
    cd comega2;
};
// psv_omega_scalars: everything but the band pieces (lsuf, calcoef,
// calabnum, calupfluid); psv_omega_band: the band pieces P0..Q1 and calbc
// from those scalars; psv_omega_setup: both.
void psv_omega_scalars(const PsvSetup &s, int i, PsvOmega &o);
void psv_omega_band(const PsvSetup &s, PsvOmega &o);
void psv_omega_setup(const PsvSetup &s, int i, PsvOmega &o);

// Frequency-independent tables of the device band assembly
// (psv_band_kernel, the operations of psv_omega_band): the element
// integrals of every zone in one array, per zone its base offset there,
// kind, layer count and first band column, per band column the (up to
// two) zones contributing to it in zone order, and calbc's overrides.
// Layout of a solid zone's block: PSV_FE_NSOLID4 arrays of 4 nlayer
// (PsvFeSolid order), then PSV_FE_NSOLIDM arrays of 4 (nlayer + 1)
// (PsvFeSolidM order); a fluid zone: p1, p2, p3, 4 nlayer each.
enum PsvFeSolid {
    FE_T, FE_H1X, FE_H2N, FE_H3AY, FE_H5AY, FE_H4AN, FE_H6AN, FE_H7Z, FE_H7Y, FE_H8N, FE_H2L, FE_H4AL, FE_H6AL, FE_H8L,
    FE_H3Y, FE_H4N, FE_H6L, FE_H4L, FE_H5Y, FE_H6N, PSV_FE_NSOLID4
};
enum PsvFeSolidM { FE_H3MY, FE_H4M2N, FE_H6M2L, FE_H4M1L, FE_H5MY, FE_H6M1N, PSV_FE_NSOLIDM };
struct PsvBandTables {
    std::vector<double> fe;
    std::vector<long long> zbase;
    std::vector<int> zsolid, znl, zkkdr;
    std::vector<int> colz;     // [nn][2], -1: none
    std::vector<int> ovr_ii;   // [nn]: band row (0-based) calbc sets, -1: none
    std::vector<double> ovr;   // [nn][2]: the value of P0 there (re, im)
};
void psv_band_tables(const PsvSetup &s, PsvBandTables &t);

// Device layouts shared by the engine and the host test build:
//  band  [nn][4 band rows][P0, P1, Q0, Q1] complex (interleaved re/im),
//  kinds [nn]: per band row ii, 4 bits (kind | second term << 2) << 4*ii.
void psv_pack_band(const PsvSetup &s, const PsvOmega &o, double *band);
void psv_pack_kind(const PsvSetup &s, int *kind);

// Rows the backward substitution must record: every receiver's six rows
// ksta-1 .. ksta+4 (interpolate(2,..) reads g0(ksta-1 ..)), the surface
// rows nn-1, nn (SYNN) -- descending -- and where each consumer finds them.
struct PsvGeom {
    std::vector<int> need;       // descending, 1-based
    std::vector<int> rec_row;    // [r][6] indices into need
    int amp_i0 = 0, amp_i1 = 0;  // rows ksta(r_n)-1, ksta(r_n)
    int syn_i0 = 0, syn_i1 = 0;  // rows nn-1, nn
    int rowlo = 0;               // lowest needed row
    std::vector<double> irf;     // 1/float(r) per receiver
};
PsvGeom psv_geometry(const PsvSetup &s);

// Per band row j (0-based): radius of the row's node and the slowest wave
// speed there (S in solids, P in fluids), for the depth cut of the batched
// forms (psv_depth_kernel).
void psv_row_profile(const PsvSetup &s, std::vector<double> &r, std::vector<double> &v);

// Interpolation weights of the reference's 3-node Taylor stencil
// (others.f90 interpolate + glu): weights b with u = sum b_i g_i.
void stencil_weights(double rsta, const double rr[3], int nderiv, double w[3]);

}  // namespace dsm

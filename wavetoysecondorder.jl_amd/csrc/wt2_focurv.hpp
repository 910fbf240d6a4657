// wt2_focurv.hpp -- the first-order (Phi, Pi) system on general meshes
// (wt2_focurv.hip): descriptor and launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace wt2 {

struct FcDesc {
    int N;
    int64_t npts;
    const double* G;         // device: differentiation matrix G[i*N+p]
    const double* w;         // device: GLL weights
    const double* xyz;       // metric set-up: node coordinates (3,N,N,N,Ne)
    const int2* conn;        // per (e, f): (neighbour or -1, fn | ori << 3 | tag << 6)
    double* JAp;             // 9 fields: sign(J) J dxi^a/dx_m at [(3a+m) npts + node]
    double* Jp;              // |J|
    double* ell;             // 3 per element: extent along each reference axis
    const double* Phi; const double* Pi;
    const double* coef;      // 11 fields (make_coef3d order)
    double* D;               // 3 fields: grad Phi
    double* F;               // 3 fields: flux
    double* fb;              // per face node (6 N^2 per element): the face-normal
                             // flux sum_m J a^f_m F^m, read by the neighbour
    double* dPhi; double* dPi;
    double* ko;              // 12 fields: the KO channels, ping-pong
    double koT;              // eps mu^-5 (0: no KO)
    int has_bc, kind, excision_tag;
    double sigma;
    const double* gPhi; const double* gPi; const double* gD;
    double* partial;         // RED_BLOCKS partial sums (energy, norms)
    int grad_only;
};

// what: 0 = metric terms; 1 = the right-hand side; 2 = the energy's partial sums
hipError_t launch_fc(const FcDesc& d, int what, hipStream_t s);
// mu of the KO derivative along axis 1 (30 power iterations on D^2 from the
// alternating start vector); v, w: two scratch fields
hipError_t fc_mu(const FcDesc& d, double* v, double* w, hipStream_t s, double* mu);

}  // namespace wt2

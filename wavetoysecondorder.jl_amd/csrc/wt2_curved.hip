// wt2_curved.hip -- curved operator (placeholder until the element kernels land).
#include <cmath>

#include "wt2_curved.hpp"
#include "wt2_internal.hpp"

namespace wt2 {

void mesh_coords(const wt2_mesh_desc& d, const double* xi_n, double* xyz)
{
    const int N = d.N;
    const int* M = d.M;
    const int64_t Ne = int64_t(M[0]) * M[1] * M[2];
    for (int64_t e = 0; e < Ne; ++e) {
        const int ec[3] = {int(e % M[0]), int((e / M[0]) % M[1]), int(e / (int64_t(M[0]) * M[1]))};
        for (int k = 0; k < N; ++k)
            for (int j = 0; j < N; ++j)
                for (int i = 0; i < N; ++i) {
                    const int ii[3] = {i, j, k};
                    double xi[3];
                    for (int a = 0; a < 3; ++a) xi[a] = (ec[a] + xi_n[ii[a]]) / M[a];
                    double s = 0.0;
                    if (d.warp_A != 0.0) {
                        const double tp = 2.0 * M_PI;
                        s = d.warp_A * std::sin(tp * xi[0]) * std::sin(tp * xi[1]) * std::sin(tp * xi[2]);
                    }
                    const int64_t n = i + int64_t(N) * (j + int64_t(N) * (k + int64_t(N) * e));
                    for (int a = 0; a < 3; ++a) xyz[3 * n + a] = d.x0[a] + (d.x1[a] - d.x0[a]) * (xi[a] + s);
                }
    }
}

int curved_init(CurvedData&, const wt2_mesh_desc&, const double*, const double*, const double*,
                const double*, hipStream_t, int64_t*)
{
    set_error("curved meshes / backgrounds are not built yet");
    return WT2_ERR_UNSUPPORTED;
}

int curved_set_background(CurvedData&, const wt2_mesh_desc&, const double*, const double*,
                          const double*, const double*, hipStream_t)
{
    set_error("curved meshes / backgrounds are not built yet");
    return WT2_ERR_UNSUPPORTED;
}

void curved_free(CurvedData& cd)
{
    hipFree(cd.coef64); hipFree(cd.coef32); hipFree(cd.wu); hipFree(cd.wv); hipFree(cd.fbuf);
    cd = CurvedData();
}

hipError_t curved_operator(const CurvedData&, const CurvedCall&, hipStream_t)
{
    return hipErrorNotSupported;
}

}  // namespace wt2

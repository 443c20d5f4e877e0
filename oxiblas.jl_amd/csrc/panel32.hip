// FP32 instances of the panel leaf (the working precision of the
// mixed-precision LU, as the reference's mixed_solve factors in Float32:
// src/lapack/extended_lapack.jl:183,195-198).  FP32 rows take half the
// registers, so up to four rows per thread stay in registers.
#include "panel_kernel.hpp"

int oxi_launch_lu_leaf32(oxi_handle_t h, hipStream_t s, const LeafArgs& a, int kind) {
    if (kind == 0) hipLaunchKernelGGL((lu_leaf_kernel<float, 256, 1>), dim3(unsigned(a.P)), dim3(256), 0, s, a);
    else if (kind == 1) hipLaunchKernelGGL((lu_leaf_kernel<float, 256, 2>), dim3(unsigned(a.P)), dim3(256), 0, s, a);
    else hipLaunchKernelGGL((lu_leaf_kernel<float, 256, 4>), dim3(unsigned(a.P)), dim3(256), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return oxi_set_error(h, OXI_ERR_HIP, "lu_leaf32 launch: %s", hipGetErrorString(e));
    // the leaf's interchanges on its stored multipliers and on the rest of its panel
    const unsigned nblk = unsigned(std::max(a.w - 1, 0) + cdiv(a.c0 - a.xlo, 64) + cdiv(a.xhi - a.c0 - a.w, 64));
    if (nblk > 0) {
        hipLaunchKernelGGL(leaf_swaps_kernel<float>, dim3(nblk), dim3(128), 0, s, (float*)a.Wm, a.ldw, a.c0, a.w,
                           (const i64*)a.ipiv, (const int*)a.flags, a.xlo, a.xhi);
        e = hipGetLastError();
        if (e != hipSuccess) return oxi_set_error(h, OXI_ERR_HIP, "leaf_swaps launch: %s", hipGetErrorString(e));
    }
    return OXI_OK;
}

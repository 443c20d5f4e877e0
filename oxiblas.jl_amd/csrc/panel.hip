// Panel leaf of the blocked LU: column-by-column elimination of an m x w
// block (w <= 32) held in registers across P co-resident workgroups.
//
// Restates the leaf of _lu_panel! (src/lapack/factorizations.jl:116-140) with
// the pivot choice of _lu_pivot (:89-107): the pivot is the first row of
// largest |value| (strict '>' scan semantics: NaN never beats a number except
// in the diagonal position), rows are swapped across the whole leaf, the
// column is divided by the pivot (a division, as the reference) and the
// leaf's remaining columns get fma(-l, u, a); a zero pivot records info (the
// first one, 1-based) and skips the column.
//
// Operands live in the transposed working layout W = A^T (see getrf.hip):
// the w leaf values of one row of A are contiguous.  Each workgroup owns a
// contiguous block of rows (one or two per thread, in registers).  Per column
// every workgroup publishes its candidate (key, row and the whole row) as a
// self-validating record (tag + checksum, panel_kernel.hpp) and reads all
// the others; every workgroup picks the same winner.  The wait is bounded:
// on timeout a flag is raised and every wave leaves, so the grid always
// drains.  With P == 1 the exchange stays in LDS.
//
// Row interchanges act on the registers (columns j and right of it) at once;
// the multipliers left of column j are already in memory and receive the
// interchanges of their later steps from a second, small kernel, which gives
// the same rows as swapping whole rows at each step.  The same kernel applies
// the leaf's interchanges to the other columns of the enclosing panel (the
// sibling swaps of the getrf2 recursion, done once per leaf).
#include "panel_kernel.hpp"

int oxi_launch_lu_leaf(oxi_handle_t h, hipStream_t s, const LeafArgs& a, int kind) {
    if (kind == 0) hipLaunchKernelGGL((lu_leaf_kernel<double, 256, 1>), dim3(unsigned(a.P)), dim3(256), 0, s, a);
    else hipLaunchKernelGGL((lu_leaf_kernel<double, 256, 2>), dim3(unsigned(a.P)), dim3(256), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return oxi_set_error(h, OXI_ERR_HIP, "lu_leaf launch: %s", hipGetErrorString(e));
    // the leaf's interchanges on its stored multipliers and on the rest of its panel
    const unsigned nblk = unsigned(std::max(a.w - 1, 0) + cdiv(a.c0 - a.xlo, 64) + cdiv(a.xhi - a.c0 - a.w, 64));
    if (nblk > 0) {
        hipLaunchKernelGGL(leaf_swaps_kernel<double>, dim3(nblk), dim3(128), 0, s, (double*)a.Wm, a.ldw, a.c0, a.w,
                           (const i64*)a.ipiv, (const int*)a.flags, a.xlo, a.xhi);
        e = hipGetLastError();
        if (e != hipSuccess) return oxi_set_error(h, OXI_ERR_HIP, "leaf_swaps launch: %s", hipGetErrorString(e));
    }
    return OXI_OK;
}

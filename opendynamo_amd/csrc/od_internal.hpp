// Internal interface between the C-ABI front end (od_api.hip) and the
// gfx950 kernels (od_kernels.hip).
//
// Device data layout (all arrays in HBM, element e = ex + nx*ey + nx*ny*ez):
//   state   [5][E*64]        UT   var-major, pt = i + 4j + 16k (file layout)
//   traces  [E][8][5][16]    CT   face traces of the state, per element:
//            0 xL, 1 xR, 2 yL, 3 yR  pairwise (t0+t1)+(t2+t3)
//            4 zL, 5 zR              sequential from 0
//            6 xL, 7 xR              sequential from the first product
//           (face point fp: x j+4k, y i+4k, z i+4j)
//   fv      [E][640]         CT   per element: [6][16][4] viscous flux
//                                 components 1..4 at the face points of the
//                                 element's own faces xL,xR,yL,yR,zL,zR, then
//                                 [64][4] the volume viscous-flux divergence of
//                                 components 1..4 at the solution points (fast
//                                 f32 build only); component fastest.  The
//                                 fast f32 kernels on 2x2x2-element blocks
//                                 store, at each face interior to a block,
//                                 the half-difference (F_vL - F_vR) / 2 in
//                                 the left element's slot only
#pragma once
#include <cstddef>
#include <cstdint>

namespace od {

constexpr int kTraceFaces = 8;
constexpr int kTraceStride = kTraceFaces * 5 * 16;  // per element
constexpr int kFvFaces = 6 * 4 * 16;                // face part, per element
constexpr int kFvStride = kFvFaces + 4 * 64;        // + volume divergence part

template <class C>
struct Ops {
    C D[16];   // D[i][m] = l_m'(x_i)
    C eL[4];   // l_i(-1)
    C eR[4];   // l_i(+1)
    C gL[4];   // g_left'(x_i)
    C gR[4];   // g_right'(x_i)
};

template <class UT, class CT>
struct StageArgs {
    const UT* uin;     // state the stage differentiates
    const UT* uprev;   // u^n (stage 1: == uin)
    UT* uout;          // updated state (may alias uin or uprev per element)
    const CT* tr_in;   // traces of uin
    CT* tr_out;        // traces of uout
    CT* fv;            // face viscous fluxes of uin
    CT* dbg;           // mode 1: corrected RHS output [5][N]
    const CT* mu_t;    // per-element eddy viscosity added to mu (null: plain FR)
    int32_t nx, ny, nz;
    int32_t wall, flux, sutherland, round_q, round_rhs, accum_state, stage1, mode;
    CT jx, jy, jz;
    CT mu, mu_ref, t_ref, s_suth, bf;
    CT ca, cb, cc;     // blend coefficients in the compute type
    UT ua, ub, uc;     // blend coefficients in the state type
};

// Copy the FR operators to the kernels' constant memory on the current
// device (both builds hold their own copy).
int upload_ops(const Ops<float>& f32, const Ops<double>& f64);

// One box of an ensemble batch.  The batch arrays hold the boxes back to
// back: box b's elements are e_off .. e_off + E_b - 1 of the trace and fv
// arrays, and its state [5][E_b][64] starts at 320 * e_off of each state
// buffer (the per-box file layout, unchanged).
template <class UT, class CT>
struct BoxDev {
    int64_t e_off;     // first element of the box
    int32_t tile_off;  // first tile of the box in the batch tile numbering
    int32_t nx, ny, nz;
    int32_t shell;     // 1: the box's core runs on 2x2x2 blocks, its tiles cover the rest (the shell)
    int32_t blk_off;   // first block of the box's core in the batch block numbering
    CT jx, jy, jz;     // 2 / (L / n) in the compute type
    CT cc[3];          // per-stage c * dt of the box in the compute type
    UT uc[3];          // and in the state type
};

constexpr int kInlineBoxes = 64;  // batches up to this size search the tile offsets in kernel arguments

template <class UT, class CT>
struct BatchArgs {
    const BoxDev<UT, CT>* boxes;  // device array, sorted by tile_off
    int32_t n_boxes, n_tiles, n_units;
    int32_t orient;               // element pair direction of a tile: 0 x, 1 y, 2 z
    int32_t order;                // 0 round robin, 1 contiguous range per XCD
    int32_t stage;                // 0..2
    int32_t tile_off[kInlineBoxes];  // first tile of each box (n_boxes <= kInlineBoxes)
};

// The boxes of a batch whose core runs on 2x2x2-element blocks (fast f32
// build; the core is the box less the last layer along each odd dimension):
// blk_off[i] = the first block of box i in the batch block numbering (a box
// without blocks: the next box's first block), held in the kernel arguments
// for batches of up to kInlineBoxes boxes and in the box table always.
struct BlockBatch {
    const BoxDev<float, float>* boxes;
    int32_t n_boxes, n_blocks, stage, pad;
    int32_t blk_off[kInlineBoxes];
};

// Kernel launchers (explicitly instantiated for UT, CT in {float, double}).
template <class UT, class CT>
int launch_trace(const StageArgs<UT, CT>& a, int64_t n_elem, void* stream);
template <class UT, class CT>
int launch_visc(const StageArgs<UT, CT>& a, int64_t n_elem, void* stream);
template <class UT, class CT, bool VISC>
int launch_update(const StageArgs<UT, CT>& a, int64_t n_elem, void* stream);
template <class UT, class CT>
int launch_batch_visc(const StageArgs<UT, CT>& a, const BatchArgs<UT, CT>& b, int pairs, int groups, void* stream);
template <class UT, class CT, bool VISC>
int launch_batch_update(const StageArgs<UT, CT>& a, const BatchArgs<UT, CT>& b, int pairs, int groups, void* stream);
// Resident work-groups per CU of a batch kernel (kind 0 visc, 1 update).
template <class UT, class CT>
int batch_occupancy(int kind, int pairs, bool visc, int* groups_per_cu);
template <class UT>
int launch_sums(const UT* u, double* partial, double* out, int64_t n_elem, const double w3[64],
                double jac, void* stream);

const char* kernel_name(int kind);

// The same kernels built with fast math (f32 compute only): reciprocal
// multiplies, hardware square root, contraction (od_kernels_fast.hip).
namespace fastmath {
template <class UT, class CT>
int launch_trace(const StageArgs<UT, CT>& a, int64_t n_elem, void* stream);
template <class UT, class CT>
int launch_visc(const StageArgs<UT, CT>& a, int64_t n_elem, void* stream);
template <class UT, class CT, bool VISC>
int launch_update(const StageArgs<UT, CT>& a, int64_t n_elem, void* stream);
template <class UT, class CT>
int launch_batch_visc(const StageArgs<UT, CT>& a, const BatchArgs<UT, CT>& b, int pairs, int groups, void* stream);
template <class UT, class CT, bool VISC>
int launch_batch_update(const StageArgs<UT, CT>& a, const BatchArgs<UT, CT>& b, int pairs, int groups, void* stream);
template <class UT, class CT>
int batch_occupancy(int kind, int pairs, bool visc, int* groups_per_cu);
int upload_ops(const Ops<float>& f32, const Ops<double>& f64);
// Kernels on blocks of 2x2x2 elements (f32 state and compute, nx, ny, nz
// even): the same results as launch_update / launch_visc, bit for bit.
// ext_traces: the update kernel writes the traces of the blocks' external
// faces only, which is all the block visc kernel reads.
template <bool VISC>
int launch_update_block(const StageArgs<float, float>& a, bool ext_traces, void* stream);
int launch_visc_block(const StageArgs<float, float>& a, void* stream);
// The same kernels over every block box of a batch, one launch each.
template <bool VISC>
int launch_batch_update_block(const StageArgs<float, float>& a, const BlockBatch& bb, void* stream);
int launch_batch_visc_block(const StageArgs<float, float>& a, const BlockBatch& bb, void* stream);
}  // namespace fastmath

}  // namespace od

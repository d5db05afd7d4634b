// Ensemble front end (od_batch_* in include/opendynamo.h).  No reference
// counterpart: drunkseal/opendynamo advances one box per process.
//
// Device layout: the boxes lie back to back in one set of buffers.  Box b
// owns elements e_off[b] .. e_off[b] + E_b - 1 of the trace and fv arrays,
// and its state [5][E_b][64] (the file layout of that box) starts at
// 320 * e_off[b] values into each of the two state buffers.  A device table
// (od::BoxDev) gives each box's dimensions, Jacobians, first tile and the
// per-stage c * dt of its own time step, so one launch per kernel and stage
// serves every box, and changing a dt only rewrites the table (the captured
// step graph stays valid).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "od_host.hpp"

using namespace od::host;

struct od_batch {
    od_params p;
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<od_box> boxes;
    std::vector<int64_t> e_off;
    int64_t E = 0, max_box_E = 0;
    size_t ubytes = 0, cbytes = 0;
    void* u[2] = {nullptr, nullptr};
    void* tr[3] = {nullptr, nullptr, nullptr};
    void* fv = nullptr;
    void* table = nullptr;
    size_t table_bytes = 0;
    double* sums_partial = nullptr;
    double* sums_out = nullptr;
    HostOps ops;
    std::vector<char> state_set;
    od_schedule sched{0, 1, 0, 0};
    std::vector<double> table_dt;  // the dt values the device table holds
    int table_orient = -1;
    int64_t n_tiles[3] = {0, 0, 0};  // tiles (boxes on tiles, shells of the others), per orientation
    int64_t n_blocks = 0;            // 2x2x2-element blocks of the boxes' cores
    bool has_shell = false;          // some box has a block core and a tile shell
    bool core_faces_stale = false;   // tr[0] lacks the block cores' internal faces (see before_steps)
    hipGraphExec_t graph = nullptr;
    int cu_count = 0;
    size_t device_bytes = 0;
};

namespace {

int64_t box_tiles(const od_box& x, int orient) {
    const int64_t h = 2;  // elements per tile
    const int64_t n0 = orient == 0 ? (x.nx + h - 1) / h : x.nx;
    const int64_t n1 = orient == 1 ? (x.ny + h - 1) / h : x.ny;
    const int64_t n2 = orient == 2 ? (x.nz + h - 1) / h : x.nz;
    return n0 * n1 * n2;
}

bool f32_compute(const od_batch* b) { return b->p.compute_real == OD_F32; }

// In fast f32 batches the core of every box at least 2 elements wide in each
// direction runs on 2x2x2-element blocks: the whole box when its dimensions are all even, else the
// box less its last layer along each odd dimension, whose elements (the
// shell) run on two-element tiles.  Blocks and tiles give the same bits for
// every face, so every box still equals the same box run alone (on blocks
// when all its dimensions are even, od_create's rule; on tiles otherwise).
bool core_box(const od_batch* b, const od_box& x) {
    return b->p.fast_math && b->p.u_real == OD_F32 && b->p.compute_real == OD_F32 && x.nx >= 2 && x.ny >= 2 &&
           x.nz >= 2;
}
bool even_box(const od_box& x) { return x.nx % 2 == 0 && x.ny % 2 == 0 && x.nz % 2 == 0; }
// Tiles of the shell (shell_tile's numbering): the x slab paired along y,
// the rest of the y and z slabs paired along x.
int64_t shell_tiles(const od_box& x) {
    const int64_t hx = x.nx / 2, hy = (x.ny + 1) / 2, cy = x.ny & ~1;
    return (x.nx % 2 ? hy * x.nz : 0) + (x.ny % 2 ? hx * x.nz : 0) + (x.nz % 2 ? hx * cy : 0);
}
int64_t tile_count(const od_batch* b, const od_box& x, int orient) {
    return core_box(b, x) ? shell_tiles(x) : box_tiles(x, orient);
}
int64_t block_count(const od_batch* b, const od_box& x) {
    return core_box(b, x) ? (int64_t)(x.nx / 2) * (x.ny / 2) * (x.nz / 2) : 0;
}

// ---- per-box launches with the single-box kernels ---------------------------
template <class UT, class CT>
od::StageArgs<UT, CT> box_args(od_batch* b, int i, int stage, double dt) {
    const int64_t eo = b->e_off[i];
    void* u[2] = {(UT*)b->u[0] + 320 * eo, (UT*)b->u[1] + 320 * eo};
    void* tr[3] = {(CT*)b->tr[0] + eo * od::kTraceStride, (CT*)b->tr[1] + eo * od::kTraceStride,
                   (CT*)b->tr[2] + eo * od::kTraceStride};
    void* fv = b->fv ? (void*)((CT*)b->fv + eo * od::kFvStride) : nullptr;
    const od_box& x = b->boxes[i];
    return make_stage_args<UT, CT>(b->p, x.nx, x.ny, x.nz, x.lx, x.ly, x.lz, u, tr, fv, nullptr, stage, dt);
}

// kind: 0 = visc, 1 = update, 2 = trace of u0 into tr[0]
template <class UT, class CT>
int box_launch_t(od_batch* b, int i, int kind, int stage, double dt) {
    od::StageArgs<UT, CT> a = box_args<UT, CT>(b, i, stage, dt);
    const int64_t E = (int64_t)b->boxes[i].nx * b->boxes[i].ny * b->boxes[i].nz;
    if (kind == 2) {
        a.uin = a.uprev;  // u0 of the box
        a.tr_out = (CT*)a.tr_in;  // tr[0] of the box (stage 0 reads tr[0])
        // the fast build's traces (one x order, fused arithmetic) come from its own kernel
        if constexpr (sizeof(CT) == 4)
            if (b->p.fast_math) return od::fastmath::launch_trace<UT, CT>(a, E, b->stream);
        return od::launch_trace<UT, CT>(a, E, b->stream);
    }
    if constexpr (sizeof(UT) == 4 && sizeof(CT) == 4) {
        if (core_box(b, b->boxes[i]) && even_box(b->boxes[i])) {  // as od_step runs this box alone
            if (kind == 0) return od::fastmath::launch_visc_block(a, b->stream);
            return b->p.viscous ? od::fastmath::launch_update_block<true>(a, true, b->stream)
                                : od::fastmath::launch_update_block<false>(a, true, b->stream);
        }
    }
    if constexpr (sizeof(CT) == 4) {
        if (b->p.fast_math) {
            if (kind == 0) return od::fastmath::launch_visc<UT, CT>(a, E, b->stream);
            if (b->p.viscous) return od::fastmath::launch_update<UT, CT, true>(a, E, b->stream);
            return od::fastmath::launch_update<UT, CT, false>(a, E, b->stream);
        }
    }
    if (kind == 0) return od::launch_visc<UT, CT>(a, E, b->stream);
    if (b->p.viscous) return od::launch_update<UT, CT, true>(a, E, b->stream);
    return od::launch_update<UT, CT, false>(a, E, b->stream);
}

int box_launch(od_batch* b, int i, int kind, int stage, double dt) {
    const bool uf = b->p.u_real == OD_F32, cf = f32_compute(b);
    if (uf && cf) return box_launch_t<float, float>(b, i, kind, stage, dt);
    if (!uf && !cf) return box_launch_t<double, double>(b, i, kind, stage, dt);
    if (!uf && cf) return box_launch_t<double, float>(b, i, kind, stage, dt);
    return box_launch_t<float, double>(b, i, kind, stage, dt);
}

// ---- the box table ---------------------------------------------------------
template <class UT, class CT>
od_status upload_table_t(od_batch* b, const double* dt, int orient) {
    std::vector<od::BoxDev<UT, CT>> t(b->boxes.size());
    int64_t tiles = 0, blocks = 0;
    for (size_t i = 0; i < b->boxes.size(); ++i) {
        const od_box& x = b->boxes[i];
        od::BoxDev<UT, CT>& d = t[i];
        std::memset(&d, 0, sizeof(d));
        d.e_off = b->e_off[i];
        d.tile_off = (int32_t)tiles;
        d.nx = x.nx;
        d.ny = x.ny;
        d.nz = x.nz;
        d.shell = core_box(b, x) ? 1 : 0;
        d.blk_off = (int32_t)blocks;
        blocks += block_count(b, x);
        d.jx = jacobian<CT>(x.lx, x.nx);
        d.jy = jacobian<CT>(x.ly, x.ny);
        d.jz = jacobian<CT>(x.lz, x.nz);
        for (int s = 0; s < 3; ++s) {
            d.cc[s] = stage_dt_coef<CT>(s, dt[i]);
            d.uc[s] = stage_dt_coef<UT>(s, dt[i]);
        }
        tiles += tile_count(b, x, orient);
    }
    HIP_TRY(hipStreamSynchronize(b->stream));
    HIP_TRY(hipMemcpy(b->table, t.data(), t.size() * sizeof(t[0]), hipMemcpyHostToDevice));
    b->table_dt.assign(dt, dt + b->boxes.size());
    b->table_orient = orient;
    if (b->sched.order == 2 && b->graph) {  // per-box launches carry dt in their arguments
        (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    return OD_OK;
}

od_status ensure_table(od_batch* b, const double* dt) {
    const size_t n = b->boxes.size();
    if (b->table_orient == b->sched.orient && b->table_dt.size() == n &&
        std::equal(b->table_dt.begin(), b->table_dt.end(), dt))
        return OD_OK;
    const bool uf = b->p.u_real == OD_F32, cf = f32_compute(b);
    if (uf && cf) return upload_table_t<float, float>(b, dt, b->sched.orient);
    if (!uf && !cf) return upload_table_t<double, double>(b, dt, b->sched.orient);
    if (!uf && cf) return upload_table_t<double, float>(b, dt, b->sched.orient);
    return upload_table_t<float, double>(b, dt, b->sched.orient);
}

// ---- batched launches ------------------------------------------------------
template <class UT, class CT>
int occupancy_t(od_batch* b, int kind, int pairs) {
    int n = 0;
    int rc = -1;
    if constexpr (sizeof(CT) == 4) {
        if (b->p.fast_math) rc = od::fastmath::batch_occupancy<UT, CT>(kind, pairs, b->p.viscous != 0, &n);
    }
    if (rc < 0) rc = od::batch_occupancy<UT, CT>(kind, pairs, b->p.viscous != 0, &n);
    return rc ? 1 : std::max(n, 1);
}

template <class UT, class CT>
int batch_launch_t(od_batch* b, int kind, int stage) {
    od::StageArgs<UT, CT> a = make_stage_args<UT, CT>(b->p, 1, 1, 1, 1.0, 1.0, 1.0, b->u, b->tr, b->fv, nullptr,
                                                      stage, 0.0);
    const od_schedule& s = b->sched;
    od::BatchArgs<UT, CT> ba{};
    ba.boxes = (const od::BoxDev<UT, CT>*)b->table;
    ba.n_boxes = (int32_t)b->boxes.size();
    ba.n_tiles = (int32_t)b->n_tiles[s.orient];
    ba.n_units = (int32_t)((b->n_tiles[s.orient] + s.pairs - 1) / s.pairs);
    ba.orient = s.orient;
    ba.order = s.order;
    ba.stage = stage;
    if (b->boxes.size() <= (size_t)od::kInlineBoxes) {
        int64_t t = 0;
        for (size_t i = 0; i < b->boxes.size(); ++i) {
            ba.tile_off[i] = (int32_t)t;
            t += tile_count(b, b->boxes[i], s.orient);
        }
    }
    int64_t g = ba.n_units;
    if (s.groups_per_cu > 0) {
        const int k = std::min(s.groups_per_cu, occupancy_t<UT, CT>(b, kind, s.pairs));
        g = std::min<int64_t>(g, (int64_t)k * b->cu_count);
    }
    if (s.order == 1) g = (g + 7) / 8 * 8;  // every XCD gets the same number of work-groups
    const int groups = (int)g;
    if constexpr (sizeof(CT) == 4) {
        if (b->p.fast_math) {
            if (kind == 0) return od::fastmath::launch_batch_visc<UT, CT>(a, ba, s.pairs, groups, b->stream);
            if (b->p.viscous) return od::fastmath::launch_batch_update<UT, CT, true>(a, ba, s.pairs, groups, b->stream);
            return od::fastmath::launch_batch_update<UT, CT, false>(a, ba, s.pairs, groups, b->stream);
        }
    }
    if (kind == 0) return od::launch_batch_visc<UT, CT>(a, ba, s.pairs, groups, b->stream);
    if (b->p.viscous) return od::launch_batch_update<UT, CT, true>(a, ba, s.pairs, groups, b->stream);
    return od::launch_batch_update<UT, CT, false>(a, ba, s.pairs, groups, b->stream);
}

// The block cores' launch of phase `kind` (0 visc, 1 update), all boxes in one.
int block_launch(od_batch* b, int kind, int stage) {
    od::StageArgs<float, float> a = make_stage_args<float, float>(b->p, 1, 1, 1, 1.0, 1.0, 1.0, b->u, b->tr, b->fv,
                                                                  nullptr, stage, 0.0);
    od::BlockBatch bb{};
    bb.boxes = (const od::BoxDev<float, float>*)b->table;
    bb.n_boxes = (int32_t)b->boxes.size();
    bb.n_blocks = (int32_t)b->n_blocks;
    bb.stage = stage;
    int64_t t = 0;
    for (size_t i = 0; i < b->boxes.size() && i < (size_t)od::kInlineBoxes; ++i) {
        bb.blk_off[i] = (int32_t)t;
        t += block_count(b, b->boxes[i]);
    }
    if (kind == 0) return od::fastmath::launch_batch_visc_block(a, bb, b->stream);
    return b->p.viscous ? od::fastmath::launch_batch_update_block<true>(a, bb, b->stream)
                        : od::fastmath::launch_batch_update_block<false>(a, bb, b->stream);
}

int batch_launch(od_batch* b, int kind, int stage) {
    const bool uf = b->p.u_real == OD_F32, cf = f32_compute(b);
    if (uf && cf) return batch_launch_t<float, float>(b, kind, stage);
    if (!uf && !cf) return batch_launch_t<double, double>(b, kind, stage);
    if (!uf && cf) return batch_launch_t<double, float>(b, kind, stage);
    return batch_launch_t<float, double>(b, kind, stage);
}

int max_occupancy(od_batch* b, int pairs) {
    const bool uf = b->p.u_real == OD_F32, cf = f32_compute(b);
    if (uf && cf) return occupancy_t<float, float>(b, 1, pairs);
    if (!uf && !cf) return occupancy_t<double, double>(b, 1, pairs);
    if (!uf && cf) return occupancy_t<double, float>(b, 1, pairs);
    return occupancy_t<float, double>(b, 1, pairs);
}

od_status enqueue_step(od_batch* b) {
    if (b->sched.order == 2) {  // box by box with the single-box kernels
        for (int s = 0; s < 3; ++s)
            for (size_t i = 0; i < b->boxes.size(); ++i) {
                const double dt = b->table_dt[i];
                if (b->p.viscous && box_launch(b, (int)i, 0, s, dt)) return fail(OD_ERR_HIP, "od_visc_kernel launch failed");
                if (box_launch(b, (int)i, 1, s, dt)) return fail(OD_ERR_HIP, "od_update_kernel launch failed");
            }
        return OD_OK;
    }
    const bool tiles = b->n_tiles[b->sched.orient] > 0, blocks = b->n_blocks > 0;
    for (int s = 0; s < 3; ++s) {
        if (b->p.viscous && tiles && batch_launch(b, 0, s))
            return fail(OD_ERR_HIP, "od_batch_visc_kernel launch failed");
        if (b->p.viscous && blocks && block_launch(b, 0, s))
            return fail(OD_ERR_HIP, "od_batch_visc_block_kernel launch failed");
        // updates walk their blocks and tiles backwards: the block cores' visc
        // output is the most recent in the memory-side cache, so their update
        // goes first
        if (blocks && block_launch(b, 1, s)) return fail(OD_ERR_HIP, "od_batch_update_block_kernel launch failed");
        if (tiles && batch_launch(b, 1, s)) return fail(OD_ERR_HIP, "od_batch_update_kernel launch failed");
    }
    return OD_OK;
}

void drop_graph(od_batch* b) {
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
}

od_status ensure_graph(od_batch* b) {
    if (b->graph) return OD_OK;
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    od_status s = enqueue_step(b);
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (s != OD_OK) {
        if (g) (void)hipGraphDestroy(g);
        return s;
    }
    if (e != hipSuccess) return fail(OD_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(OD_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    e = hipGraphUpload(b->graph, b->stream);  // device-side set-up now, not in the first replay
    if (e != hipSuccess) return fail(OD_ERR_HIP, std::string("hipGraphUpload: ") + hipGetErrorString(e));
    return OD_OK;
}

od_status check_dt(const od_batch* b, const double* dt) {
    if (!dt) return fail(OD_ERR_INVALID, "dt: null argument");
    for (size_t i = 0; i < b->boxes.size(); ++i)
        if (!(dt[i] > 0.0) || !std::isfinite(dt[i])) return fail(OD_ERR_INVALID, "dt must be positive and finite");
    return OD_OK;
}

od_status ready(od_batch* b, const char* who) {
    for (char c : b->state_set)
        if (!c) return fail(OD_ERR_STATE, std::string(who) + ": the state of every box must be set first");
    return OD_OK;
}

od_status check_schedule(const od_schedule* s) {
    if (!s) return fail(OD_ERR_INVALID, "schedule: null argument");
    if (s->orient < 0 || s->orient > 2) return fail(OD_ERR_INVALID, "schedule.orient must be 0, 1 or 2");
    if (s->pairs != 1 && s->pairs != 2) return fail(OD_ERR_INVALID, "schedule.pairs must be 1 or 2");
    if (s->order < 0 || s->order > 2) return fail(OD_ERR_INVALID, "schedule.order must be 0, 1 or 2");
    if (s->groups_per_cu < 0 || s->groups_per_cu > 64) return fail(OD_ERR_INVALID, "schedule.groups_per_cu must be 0..64");
    return OD_OK;
}

od_status before_steps(od_batch* b, bool per_box);

od_status time_graph(od_batch* b, uint32_t n_steps, double* ms) {
    od_status st = before_steps(b, b->sched.order == 2);
    if (st != OD_OK) return st;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, b->stream));
    for (uint32_t q = 0; q < n_steps; ++q) HIP_TRY(hipGraphLaunch(b->graph, b->stream));
    HIP_TRY(hipEventRecord(e1, b->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float f = 0.f;
    HIP_TRY(hipEventElapsedTime(&f, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = f;
    return OD_OK;
}

// Traces tr[0] of every box from its u0 (after a state change).
od_status retrace_all(od_batch* b) {
    for (size_t i = 0; i < b->boxes.size(); ++i)
        if (box_launch(b, (int)i, 2, 0, 0.0)) return fail(OD_ERR_HIP, "od_trace_kernel launch failed");
    b->core_faces_stale = false;
    return OD_OK;
}

// Before steps on the batch's stream.  Per-box launches (schedule order 2,
// per-box timing) run a box with a shell on tiles alone, and tiles read the
// traces of every face of tr[0]; batched steps leave the block cores'
// internal faces out of it (the block kernels form those from the state), so
// after batched steps they are formed again from u0 first.  The trace kernel
// forms them with the arithmetic of the update kernels' trace stores.
od_status before_steps(od_batch* b, bool per_box) {
    if (per_box && b->core_faces_stale) return retrace_all(b);
    if (!per_box && b->has_shell) b->core_faces_stale = true;
    return OD_OK;
}

}  // namespace

extern "C" {

od_status od_batch_create(const od_params* p, const od_box* boxes, int32_t n_boxes, int device, od_batch** out) {
    if (!p || !boxes || !out) return fail(OD_ERR_INVALID, "od_batch_create: null argument");
    *out = nullptr;
    if (n_boxes < 1) return fail(OD_ERR_INVALID, "od_batch_create: at least one box is needed");
    if (!real_ok(p->u_real) || !real_ok(p->q_real) || !real_ok(p->rhs_real) || !real_ok(p->compute_real))
        return fail(OD_ERR_INVALID, "precision types must be OD_F32 or OD_F64");
    if (p->flux_scheme < 0 || p->flux_scheme > 2) return fail(OD_ERR_INVALID, "flux_scheme must be 0..2");
    if (p->viscosity_model < 0 || p->viscosity_model > 1) return fail(OD_ERR_INVALID, "viscosity_model must be 0 or 1");
    int64_t E = 0, tiles_max = 0, max_box = 0;
    for (int i = 0; i < n_boxes; ++i) {
        const od_box& x = boxes[i];
        if (x.nx < 1 || x.ny < 1 || x.nz < 1) return fail(OD_ERR_INVALID, "box dimensions must be >= 1");
        if (x.ny > 65535 || x.nz > 65535) return fail(OD_ERR_INVALID, "box ny and nz must be <= 65535");
        if (!(x.lx > 0 && x.ly > 0 && x.lz > 0)) return fail(OD_ERR_INVALID, "box extents must be positive");
        const int64_t e = (int64_t)x.nx * x.ny * x.nz;
        if (e > (int64_t)0x7fffffff) return fail(OD_ERR_INVALID, "too many elements in one box");
        E += e;
        max_box = std::max(max_box, e);
    }
    for (int o = 0; o < 3; ++o) {
        int64_t t = 0;
        for (int i = 0; i < n_boxes; ++i) t += box_tiles(boxes[i], o);
        tiles_max = std::max(tiles_max, t);
    }
    if (tiles_max * 256 > (int64_t)0xffffffff) return fail(OD_ERR_INVALID, "batch too large for one launch grid");
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(OD_ERR_INVALID, "no such HIP device");
    HIP_TRY(hipSetDevice(device));

    od_batch* b = new od_batch();
    b->p = *p;
    b->device = device;
    b->boxes.assign(boxes, boxes + n_boxes);
    b->state_set.assign(n_boxes, 0);
    b->E = E;
    b->max_box_E = max_box;
    b->ubytes = (size_t)p->u_real;
    b->cbytes = (size_t)p->compute_real;
    b->ops = fr_operators();
    auto cleanup = [&](od_status s) {
        od_batch_destroy(b);
        return s;
    };
    {
        const od::Ops<float> of = cast_ops<float>(b->ops);
        const od::Ops<double> o64 = cast_ops<double>(b->ops);
        if (od::upload_ops(of, o64) || od::fastmath::upload_ops(of, o64))
            return cleanup(fail(OD_ERR_HIP, "copying the FR operators to the device failed"));
    }
    int64_t off = 0;
    for (int i = 0; i < n_boxes; ++i) {
        b->e_off.push_back(off);
        off += (int64_t)boxes[i].nx * boxes[i].ny * boxes[i].nz;
    }
    for (int o = 0; o < 3; ++o)
        for (int i = 0; i < n_boxes; ++i) b->n_tiles[o] += tile_count(b, boxes[i], o);
    for (int i = 0; i < n_boxes; ++i) {
        b->n_blocks += block_count(b, boxes[i]);
        b->has_shell = b->has_shell || (core_box(b, boxes[i]) && !even_box(boxes[i]));
    }
    if (b->n_blocks * 512 > (int64_t)0xffffffff)
        return cleanup(fail(OD_ERR_INVALID, "batch too large for one launch grid"));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    b->cu_count = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(OD_ERR_HIP, "stream creation failed"));
    auto dmalloc = [&](void** ptr, size_t bytes) {
        if (hipMalloc(ptr, bytes) != hipSuccess) return false;
        b->device_bytes += bytes;
        return true;
    };
    const size_t state_bytes = (size_t)5 * 64 * E * b->ubytes;
    bool ok = dmalloc(&b->u[0], state_bytes) && dmalloc(&b->u[1], state_bytes);
    for (int q = 0; q < 3 && ok; ++q) ok = dmalloc(&b->tr[q], (size_t)E * od::kTraceStride * b->cbytes);
    ok = ok && (!p->viscous || dmalloc(&b->fv, (size_t)E * od::kFvStride * b->cbytes));
    b->table_bytes = (size_t)n_boxes * sizeof(od::BoxDev<double, double>);
    ok = ok && dmalloc(&b->table, b->table_bytes);
    ok = ok && dmalloc((void**)&b->sums_partial, (size_t)5 * max_box * sizeof(double)) &&
         dmalloc((void**)&b->sums_out, 8 * sizeof(double));
    if (!ok) return cleanup(fail(OD_ERR_NOMEM, "device allocation failed"));
    *out = b;
    return OD_OK;
}

void od_batch_destroy(od_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    for (void* q : b->u)
        if (q) (void)hipFree(q);
    for (void* q : b->tr)
        if (q) (void)hipFree(q);
    if (b->fv) (void)hipFree(b->fv);
    if (b->table) (void)hipFree(b->table);
    if (b->sums_partial) (void)hipFree(b->sums_partial);
    if (b->sums_out) (void)hipFree(b->sums_out);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

od_status od_batch_set_state(od_batch* b, int32_t box, const void* host, size_t n, od_real dtype) {
    if (!b || !host) return fail(OD_ERR_INVALID, "od_batch_set_state: null argument");
    if (box < 0 || box >= (int32_t)b->boxes.size()) return fail(OD_ERR_INVALID, "od_batch_set_state: no such box");
    const od_box& x = b->boxes[box];
    const size_t want = (size_t)5 * 64 * x.nx * x.ny * x.nz;
    if (n != want) return fail(OD_ERR_INVALID, "od_batch_set_state: wrong number of values");
    if (!real_ok(dtype)) return fail(OD_ERR_INVALID, "od_batch_set_state: bad dtype");
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipStreamSynchronize(b->stream));
    std::vector<unsigned char> tmp;
    const void* src = host;
    if ((int)dtype != b->p.u_real) {
        tmp.resize(n * b->ubytes);
        if (b->p.u_real == OD_F64)
            convert((double*)tmp.data(), (const float*)host, n);
        else
            convert((float*)tmp.data(), (const double*)host, n);
        src = tmp.data();
    }
    unsigned char* dst = (unsigned char*)b->u[0] + (size_t)320 * b->e_off[box] * b->ubytes;
    HIP_TRY(hipMemcpyAsync(dst, src, n * b->ubytes, hipMemcpyHostToDevice, b->stream));
    if (box_launch(b, box, 2, 0, 0.0)) return fail(OD_ERR_HIP, "od_trace_kernel launch failed");
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->state_set[box] = 1;
    return OD_OK;
}

od_status od_batch_get_state(od_batch* b, int32_t box, void* host, size_t n, od_real dtype) {
    if (!b || !host) return fail(OD_ERR_INVALID, "od_batch_get_state: null argument");
    if (box < 0 || box >= (int32_t)b->boxes.size()) return fail(OD_ERR_INVALID, "od_batch_get_state: no such box");
    const od_box& x = b->boxes[box];
    if (n != (size_t)5 * 64 * x.nx * x.ny * x.nz) return fail(OD_ERR_INVALID, "od_batch_get_state: wrong number of values");
    if (!real_ok(dtype)) return fail(OD_ERR_INVALID, "od_batch_get_state: bad dtype");
    HIP_TRY(hipSetDevice(b->device));
    const unsigned char* src = (const unsigned char*)b->u[0] + (size_t)320 * b->e_off[box] * b->ubytes;
    std::vector<unsigned char> tmp;
    void* dst = host;
    if ((int)dtype != b->p.u_real) {
        tmp.resize(n * b->ubytes);
        dst = tmp.data();
    }
    HIP_TRY(hipMemcpyAsync(dst, src, n * b->ubytes, hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    if ((int)dtype != b->p.u_real) {
        if (b->p.u_real == OD_F64)
            convert((float*)host, (const double*)tmp.data(), n);
        else
            convert((double*)host, (const float*)tmp.data(), n);
    }
    return OD_OK;
}

od_status od_batch_step(od_batch* b, const double* dt, uint32_t n_steps) {
    if (!b) return fail(OD_ERR_INVALID, "od_batch_step: null batch");
    od_status s = check_dt(b, dt);
    if (s == OD_OK) s = ready(b, "od_batch_step");
    if (s != OD_OK) return s;
    HIP_TRY(hipSetDevice(b->device));
    if ((s = ensure_table(b, dt)) != OD_OK || (s = ensure_graph(b)) != OD_OK ||
        (s = before_steps(b, b->sched.order == 2)) != OD_OK)
        return s;
    for (uint32_t q = 0; q < n_steps; ++q) HIP_TRY(hipGraphLaunch(b->graph, b->stream));
    return OD_OK;
}

od_status od_batch_synchronize(od_batch* b) {
    if (!b) return fail(OD_ERR_INVALID, "od_batch_synchronize: null batch");
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return OD_OK;
}

od_status od_batch_cfl_dt(od_batch* b, double cfl, double* dt_out) {
    if (!b || !dt_out) return fail(OD_ERR_INVALID, "od_batch_cfl_dt: null argument");
    if (!(cfl > 0.0)) return fail(OD_ERR_INVALID, "od_batch_cfl_dt: cfl must be positive");
    for (size_t i = 0; i < b->boxes.size(); ++i) {
        const od_box& x = b->boxes[i];
        const int64_t NP = (int64_t)64 * x.nx * x.ny * x.nz;
        std::vector<double> h((size_t)5 * NP);
        od_status s = od_batch_get_state(b, (int32_t)i, h.data(), h.size(), OD_F64);
        if (s != OD_OK) return s;
        dt_out[i] = cfl_dt_from_state(h.data(), NP, x.lx / x.nx, x.ly / x.ny, x.lz / x.nz, cfl);
    }
    return OD_OK;
}

od_status od_batch_conservation_sums(od_batch* b, double* out) {
    if (!b || !out) return fail(OD_ERR_INVALID, "od_batch_conservation_sums: null argument");
    HIP_TRY(hipSetDevice(b->device));
    double w3[64];
    for (int pt = 0; pt < 64; ++pt) w3[pt] = b->ops.w[pt & 3] * b->ops.w[(pt >> 2) & 3] * b->ops.w[pt >> 4];
    for (size_t i = 0; i < b->boxes.size(); ++i) {
        const od_box& x = b->boxes[i];
        const int64_t E = (int64_t)x.nx * x.ny * x.nz;
        const double jac = (x.lx / x.nx) * (x.ly / x.ny) * (x.lz / x.nz) / 8.0;
        const size_t so = (size_t)320 * b->e_off[i];
        int rc = b->p.u_real == OD_F32
                     ? od::launch_sums<float>((const float*)b->u[0] + so, b->sums_partial, b->sums_out, E, w3, jac,
                                              b->stream)
                     : od::launch_sums<double>((const double*)b->u[0] + so, b->sums_partial, b->sums_out, E, w3, jac,
                                               b->stream);
        if (rc) return fail(OD_ERR_HIP, "conservation sum kernels failed to launch");
        HIP_TRY(hipMemcpyAsync(out + 5 * i, b->sums_out, 5 * sizeof(double), hipMemcpyDeviceToHost, b->stream));
        HIP_TRY(hipStreamSynchronize(b->stream));
    }
    return OD_OK;
}

od_status od_batch_set_schedule(od_batch* b, const od_schedule* s) {
    if (!b) return fail(OD_ERR_INVALID, "od_batch_set_schedule: null batch");
    od_status st = check_schedule(s);
    if (st != OD_OK) return st;
    if (std::memcmp(s, &b->sched, sizeof(*s)) == 0) return OD_OK;
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->sched = *s;
    drop_graph(b);
    return OD_OK;
}

od_status od_batch_get_schedule(od_batch* b, od_schedule* s) {
    if (!b || !s) return fail(OD_ERR_INVALID, "od_batch_get_schedule: null argument");
    *s = b->sched;
    return OD_OK;
}

od_status od_batch_autotune(od_batch* b, const double* dt, uint32_t steps_per_trial, od_schedule* best,
                            double* best_ms, od_schedule* trials, double* trial_ms, int32_t cap,
                            int32_t* n_trials) {
    if (!b) return fail(OD_ERR_INVALID, "od_batch_autotune: null batch");
    od_status st = check_dt(b, dt);
    if (st == OD_OK) st = ready(b, "od_batch_autotune");
    if (st != OD_OK) return st;
    if (steps_per_trial < 1) return fail(OD_ERR_INVALID, "od_batch_autotune: steps_per_trial must be >= 1");
    HIP_TRY(hipSetDevice(b->device));
    std::vector<od_schedule> cand;
    cand.push_back(od_schedule{0, 1, 2, 0});  // box by box
    for (int pairs = 1; pairs <= 2; ++pairs) {
        const int occ = max_occupancy(b, pairs);
        std::vector<int> gpcs = {0, 1};
        if (occ >= 2) gpcs.push_back(2);
        if (occ >= 3) gpcs.push_back(occ);
        for (int orient = 0; orient < 3; ++orient)
            for (int order = 0; order < 2; ++order)
                for (int g : gpcs) cand.push_back(od_schedule{orient, pairs, order, g});
    }
    const size_t state_bytes = (size_t)5 * 64 * b->E * b->ubytes;
    const od_schedule orig = b->sched;
    void* bak = nullptr;
    bool backed_up = false;
    if (hipMalloc(&bak, state_bytes) != hipSuccess) return fail(OD_ERR_NOMEM, "od_batch_autotune: backup allocation failed");
    // On failure the batch is left as it was: the state from before the
    // trials (and its traces) and the original schedule; the first error's
    // message is kept.
    auto done = [&](od_status s) {
        (void)hipStreamSynchronize(b->stream);
        if (s != OD_OK && backed_up) {
            const std::string msg = od_last_error();
            if (std::memcmp(&b->sched, &orig, sizeof(orig)) != 0) {
                b->sched = orig;
                drop_graph(b);
            }
            (void)hipMemcpyAsync(b->u[0], bak, state_bytes, hipMemcpyDeviceToDevice, b->stream);
            (void)retrace_all(b);
            (void)hipStreamSynchronize(b->stream);
            s = fail(s, msg);
        }
        (void)hipFree(bak);
        return s;
    };
    if (hipMemcpyAsync(bak, b->u[0], state_bytes, hipMemcpyDeviceToDevice, b->stream) != hipSuccess)
        return done(fail(OD_ERR_HIP, "od_batch_autotune: backup copy failed"));
    backed_up = true;
    od_schedule win = orig;
    double win_ms = 1e300;
    int32_t count = 0;
    for (const od_schedule& c : cand) {
        if ((st = od_batch_set_schedule(b, &c)) != OD_OK || (st = ensure_table(b, dt)) != OD_OK ||
            (st = ensure_graph(b)) != OD_OK)
            return done(st);
        double ms = 0.0;
        if ((st = before_steps(b, c.order == 2)) != OD_OK) return done(st);
        if (hipGraphLaunch(b->graph, b->stream) != hipSuccess)
            return done(fail(OD_ERR_HIP, "od_batch_autotune: graph launch failed"));
        if ((st = time_graph(b, steps_per_trial, &ms)) != OD_OK) return done(st);
        ms /= steps_per_trial;
        if (trials && trial_ms && count < cap) {
            trials[count] = c;
            trial_ms[count] = ms;
        }
        ++count;
        if (ms < win_ms) {
            win_ms = ms;
            win = c;
        }
        if (hipMemcpyAsync(b->u[0], bak, state_bytes, hipMemcpyDeviceToDevice, b->stream) != hipSuccess)
            return done(fail(OD_ERR_HIP, "od_batch_autotune: restore copy failed"));
        if ((st = retrace_all(b)) != OD_OK) return done(st);
    }
    if ((st = od_batch_set_schedule(b, &win)) != OD_OK) return done(st);
    if (best) *best = win;
    if (best_ms) *best_ms = win_ms;
    if (n_trials) *n_trials = count;
    return done(OD_OK);
}

od_status od_batch_time_steps(od_batch* b, const double* dt, uint32_t n_steps, int32_t per_box, double* ms_total) {
    if (!b || !ms_total) return fail(OD_ERR_INVALID, "od_batch_time_steps: null argument");
    od_status st = check_dt(b, dt);
    if (st == OD_OK) st = ready(b, "od_batch_time_steps");
    if (st != OD_OK) return st;
    HIP_TRY(hipSetDevice(b->device));
    if (!per_box) {
        if ((st = ensure_table(b, dt)) != OD_OK || (st = ensure_graph(b)) != OD_OK) return st;
        return time_graph(b, n_steps, ms_total);
    }
    if ((st = before_steps(b, true)) != OD_OK) return st;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, b->stream));
    for (uint32_t q = 0; q < n_steps; ++q)
        for (int s = 0; s < 3; ++s)
            for (size_t i = 0; i < b->boxes.size(); ++i) {
                if (b->p.viscous && box_launch(b, (int)i, 0, s, dt[i])) return fail(OD_ERR_HIP, "od_visc_kernel launch failed");
                if (box_launch(b, (int)i, 1, s, dt[i])) return fail(OD_ERR_HIP, "od_update_kernel launch failed");
            }
    HIP_TRY(hipEventRecord(e1, b->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float f = 0.f;
    HIP_TRY(hipEventElapsedTime(&f, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_total = f;
    return OD_OK;
}

od_status od_batch_info(od_batch* b, double out[4]) {
    if (!b || !out) return fail(OD_ERR_INVALID, "od_batch_info: null argument");
    out[0] = 64.0 * (double)b->E;
    out[1] = (double)b->device_bytes;
    const int kinds = (b->n_tiles[b->sched.orient] > 0 ? 1 : 0) + (b->n_blocks > 0 ? 1 : 0);
    out[2] = (b->p.viscous ? 6.0 : 3.0) * kinds;
    out[3] = (double)b->boxes.size();
    return OD_OK;
}

}  // extern "C"

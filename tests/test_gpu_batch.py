"""Ensembles (od_batch_*): every box of a batch advanced with one launch per
kernel form and stage must equal the same box run alone through od_step,
byte for byte, for every schedule the autotuner can pick; tuning leaves the
state unchanged.  Boxes have odd and even sizes in every direction so that
tiles with a missing second element occur for each tile orientation (the
strict, f64 and mixed variants run every box on tiles; the fast f32 ones run
each box's core on 2x2x2-element blocks and its odd layers on tiles)."""
import numpy as np
import pytest

import opendynamo_amd as od
from helpers import smooth_state
from oracle import oracle as O

pytestmark = pytest.mark.gpu

BOXES = [(3, 4, 5, 6.0, 2.0, 3.0), (4, 4, 4, 6.283185307179586, 6.283185307179586, 6.283185307179586),
         (5, 3, 2, 4.0, 2.0, 2.5), (2, 5, 3, 3.0, 5.0, 3.0)]
DT = [0.004, 0.003, 0.0025, 0.0035]

SCHEDULES = [
    dict(orient=0, pairs=1, order=0, groups_per_cu=0),
    dict(orient=1, pairs=2, order=1, groups_per_cu=1),
    dict(orient=2, pairs=1, order=1, groups_per_cu=2),
    dict(orient=0, pairs=2, order=0, groups_per_cu=1),
    dict(orient=0, pairs=1, order=2, groups_per_cu=0),  # box by box inside the graph
]

VARIANTS = {
    "f32_fast_ns_wall": dict(u_real=od.OD_F32, flux="slau2", y_wall=True, viscosity_model=1, mu_ref=0.002,
                             fast_math=1),
    "f32_strict_ns": dict(u_real=od.OD_F32, flux="ausm+", y_wall=False, mu_ref=0.002, fast_math=0),
    "f64_ns_wall": dict(u_real=od.OD_F64, flux="slau2", y_wall=True, mu_ref=0.002),
    "mixed_euler": dict(u_real=od.OD_F64, compute_real=od.OD_F32, q_real=od.OD_F32, rhs_real=od.OD_F32,
                        flux="rusanov", viscous=False, body_force_x=0.0, rk_accum=1),
    "mixed_ns_fast": dict(u_real=od.OD_F64, compute_real=od.OD_F32, q_real=od.OD_F32, rhs_real=od.OD_F32,
                          flux="slau2", y_wall=True, mu_ref=0.002, rk_accum=1, fast_math=1),
}


def announce(**case):
    print("CASE", " ".join(f"{k}={v}" for k, v in case.items()), flush=True)


def initial(q, dtype):
    nx, ny, nz, lx, ly, lz = BOXES[q]
    return smooth_state(nx, ny, nz, lx, ly, lz, seed=10 + q, amp=0.04, dtype=dtype)


def single_runs(kw, n_steps):
    out = []
    for q, (nx, ny, nz, lx, ly, lz) in enumerate(BOXES):
        p = od.params(nx, ny, nz, lx, ly, lz, **kw)
        with od.Solver(p) as s:
            s.set_state(initial(q, s.dtype))
            s.step(DT[q], n_steps)
            out.append(s.get_state())
    return out


def batch_params(kw):
    return od.params(1, 1, 1, 1.0, 1.0, 1.0, **kw)


@pytest.mark.parametrize("variant", list(VARIANTS))
def test_batch_equals_boxes_run_alone(variant):
    kw = VARIANTS[variant]
    n_steps = 3
    ref = single_runs(kw, n_steps)
    with od.Batch(batch_params(kw), BOXES) as b:
        for sched in SCHEDULES:
            announce(test="batch_equals_alone", variant=variant, **sched)
            b.set_schedule(**sched)
            for q in range(len(BOXES)):
                b.set_state(q, initial(q, b.dtype))
            b.step(DT, n_steps)
            for q in range(len(BOXES)):
                got = b.get_state(q)
                assert got.tobytes() == ref[q].tobytes(), (variant, sched, q)


def test_batch_f64_euler_matches_oracle():
    """Transitively pinned above; checked directly once against the oracle."""
    kw = dict(u_real=od.OD_F64, flux="rusanov", viscous=False, body_force_x=0.0)
    boxes = BOXES[:2]
    with od.Batch(batch_params(kw), boxes) as b:
        b.set_schedule(orient=1, pairs=2, order=1, groups_per_cu=1)
        for q in range(len(boxes)):
            b.set_state(q, initial(q, np.float64))
        b.step(DT[:2], 2)
        for q, (nx, ny, nz, lx, ly, lz) in enumerate(boxes):
            announce(test="batch_vs_oracle", box=q)
            po = O.params(nx, ny, nz, float(np.float32(lx)), float(np.float32(ly)), float(np.float32(lz)),
                          flux="rusanov", viscous=False, body_force_x=0.0)
            want = O.step(po, initial(q, np.float64), float(np.float32(DT[q])), 2)
            assert b.get_state(q).tobytes() == np.asarray(want, dtype=np.float64).tobytes()


def test_autotune_keeps_state_and_result():
    kw = VARIANTS["f32_fast_ns_wall"]
    ref = single_runs(kw, 2)
    with od.Batch(batch_params(kw), BOXES) as b:
        for q in range(len(BOXES)):
            b.set_state(q, initial(q, b.dtype))
        before = [b.get_state(q).tobytes() for q in range(len(BOXES))]
        announce(test="autotune")
        best, best_ms, table = b.autotune(DT, steps_per_trial=1)
        assert len(table) >= 24 and best_ms > 0
        assert best == b.get_schedule()
        assert min(t["ms_per_step"] for t in table) == best_ms
        assert all(t["ms_per_step"] > 0 for t in table)
        assert [b.get_state(q).tobytes() for q in range(len(BOXES))] == before
        b.step(DT, 2)
        for q in range(len(BOXES)):
            assert b.get_state(q).tobytes() == ref[q].tobytes()


def test_batch_sums_cfl_info_and_errors():
    kw = VARIANTS["f64_ns_wall"]
    with od.Batch(batch_params(kw), BOXES) as b:
        with pytest.raises(od.ODError):
            b.step(DT, 1)  # states not set
        for q in range(len(BOXES)):
            b.set_state(q, initial(q, np.float64))
        with pytest.raises(od.ODError):
            b.set_schedule(orient=3)
        with pytest.raises(od.ODError):
            b.set_state(0, np.zeros(7))
        sums = b.sums()
        dts = b.cfl_dt(0.3)
        info = b.info()
        assert info["boxes"] == len(BOXES) and info["launches_per_step"] == 6
        assert info["dof"] == 64 * sum(x[0] * x[1] * x[2] for x in BOXES)
    for q, (nx, ny, nz, lx, ly, lz) in enumerate(BOXES):
        with od.Solver(od.params(nx, ny, nz, lx, ly, lz, **kw)) as s:
            s.set_state(initial(q, np.float64))
            assert s.sums().tobytes() == sums[q].tobytes()
            assert s.cfl_dt(0.3) == dts[q]


def test_per_box_launches_equal_batched():
    kw = VARIANTS["f32_fast_ns_wall"]
    ref = single_runs(kw, 2)
    with od.Batch(batch_params(kw), BOXES) as b:
        for q in range(len(BOXES)):
            b.set_state(q, initial(q, b.dtype))
        announce(test="per_box_timing")
        ms = b.time_steps(DT, 2, per_box=True)
        assert ms > 0
        for q in range(len(BOXES)):
            assert b.get_state(q).tobytes() == ref[q].tobytes()


@pytest.mark.parametrize("order", [0, 2])
def test_changing_dt_between_steps(order):
    """A new dt rewrites the box table (batched) or re-captures the per-box
    graph; both must follow the dt sequence of the boxes run alone."""
    kw = VARIANTS["f32_fast_ns_wall"]
    dt2 = [d * 0.5 for d in DT]
    ref = []
    for q, (nx, ny, nz, lx, ly, lz) in enumerate(BOXES):
        with od.Solver(od.params(nx, ny, nz, lx, ly, lz, **kw)) as s:
            s.set_state(initial(q, s.dtype))
            s.step(DT[q], 2)
            s.step(dt2[q], 1)
            ref.append(s.get_state())
    announce(test="dt_change", order=order)
    with od.Batch(batch_params(kw), BOXES) as b:
        b.set_schedule(orient=0, pairs=1, order=order, groups_per_cu=0)
        for q in range(len(BOXES)):
            b.set_state(q, initial(q, b.dtype))
        b.step(DT, 2)
        b.step(dt2, 1)
        for q in range(len(BOXES)):
            assert b.get_state(q).tobytes() == ref[q].tobytes()


# In fast f32 batches the core of every box runs the 2x2x2-element block
# kernels and the last layer along each odd dimension (the shell) the tile
# kernels.  Even boxes (all blocks) come first, last and between boxes with a
# shell, whose odd dimensions take every combination with x: x and y, x and z,
# y and z, all three.
MIXED_BOXES = [(4, 2, 2, 4.0, 2.0, 2.0), (3, 3, 2, 3.0, 3.0, 2.0), (2, 4, 6, 2.0, 4.0, 6.0),
               (6, 2, 4, 6.0, 2.0, 4.0), (5, 2, 3, 5.0, 2.0, 3.0), (2, 3, 5, 2.0, 3.0, 5.0),
               (3, 3, 3, 3.0, 3.0, 3.0), (4, 4, 4, 4.0, 4.0, 4.0)]
MIXED_DT = [0.004, 0.003, 0.0025, 0.0035, 0.003, 0.0028, 0.0032, 0.002]
MIXED_VARIANTS = {
    "ns_wall": VARIANTS["f32_fast_ns_wall"],
    "euler": dict(u_real=od.OD_F32, flux="rusanov", viscous=False, body_force_x=0.0, fast_math=1),
}


@pytest.mark.parametrize("variant", list(MIXED_VARIANTS))
def test_block_and_tile_boxes_equal_boxes_run_alone(variant):
    kw = MIXED_VARIANTS[variant]
    n_steps = 3

    def init(q, dtype):
        nx, ny, nz, lx, ly, lz = MIXED_BOXES[q]
        return smooth_state(nx, ny, nz, lx, ly, lz, seed=40 + q, amp=0.04, dtype=dtype)

    ref = []
    for q, (nx, ny, nz, lx, ly, lz) in enumerate(MIXED_BOXES):
        with od.Solver(od.params(nx, ny, nz, lx, ly, lz, **kw)) as s:
            s.set_state(init(q, s.dtype))
            s.step(MIXED_DT[q], n_steps)
            ref.append(s.get_state())
    with od.Batch(batch_params(kw), MIXED_BOXES) as b:
        per_phase = 2 if kw.get("viscous", True) else 1
        assert b.info()["launches_per_step"] == 3 * per_phase * 2  # tile and block launches
        for sched in SCHEDULES:
            announce(test="block_and_tile_boxes", variant=variant, **sched)
            b.set_schedule(**sched)
            for q in range(len(MIXED_BOXES)):
                b.set_state(q, init(q, b.dtype))
            b.step(MIXED_DT, n_steps)
            for q in range(len(MIXED_BOXES)):
                assert b.get_state(q).tobytes() == ref[q].tobytes(), (variant, sched, q)
        announce(test="block_and_tile_boxes_per_box", variant=variant)
        for q in range(len(MIXED_BOXES)):
            b.set_state(q, init(q, b.dtype))
        b.time_steps(MIXED_DT, n_steps, per_box=True)
        for q in range(len(MIXED_BOXES)):
            assert b.get_state(q).tobytes() == ref[q].tobytes(), (variant, "per box", q)


def test_all_block_boxes():
    """A batch of even boxes only: no tile launch at all."""
    kw = VARIANTS["f32_fast_ns_wall"]
    boxes = [MIXED_BOXES[0], MIXED_BOXES[2], MIXED_BOXES[7]]
    dts = [MIXED_DT[0], MIXED_DT[2], MIXED_DT[7]]
    inits = [smooth_state(*x, seed=60 + q, amp=0.04, dtype=np.float32) for q, x in enumerate(boxes)]
    ref = []
    for q, x in enumerate(boxes):
        with od.Solver(od.params(*x, **kw)) as s:
            s.set_state(inits[q])
            s.step(dts[q], 2)
            ref.append(s.get_state())
    announce(test="all_block_boxes")
    with od.Batch(batch_params(kw), boxes) as b:
        assert b.info()["launches_per_step"] == 6
        for q in range(len(boxes)):
            b.set_state(q, inits[q])
        b.step(dts, 2)
        for q in range(len(boxes)):
            assert b.get_state(q).tobytes() == ref[q].tobytes(), q


def test_switching_between_batched_and_per_box_steps():
    """Batched steps leave the internal faces of the block cores out of the
    traces; per-box steps run a box with a shell on tiles alone, which read
    every face, so the batch forms them again first.  Steps alternate between
    the batched graph, the per-box graph and per-box launches without a new
    state in between, and every box must follow the box run alone."""
    kw = MIXED_VARIANTS["ns_wall"]
    inits = [smooth_state(*x, seed=80 + q, amp=0.04, dtype=np.float32) for q, x in enumerate(MIXED_BOXES)]
    ref = []
    for q, x in enumerate(MIXED_BOXES):
        with od.Solver(od.params(*x, **kw)) as s:
            s.set_state(inits[q])
            s.step(MIXED_DT[q], 7)
            ref.append(s.get_state())
    announce(test="switch_batched_per_box")
    with od.Batch(batch_params(kw), MIXED_BOXES) as b:
        for q in range(len(MIXED_BOXES)):
            b.set_state(q, inits[q])
        b.set_schedule(orient=0, pairs=1, order=0, groups_per_cu=0)
        b.step(MIXED_DT, 2)
        b.set_schedule(orient=0, pairs=1, order=2, groups_per_cu=0)
        b.step(MIXED_DT, 1)
        b.set_schedule(orient=2, pairs=2, order=1, groups_per_cu=0)
        b.step(MIXED_DT, 2)
        b.time_steps(MIXED_DT, 1, per_box=True)
        b.step(MIXED_DT, 1)
        for q in range(len(MIXED_BOXES)):
            assert b.get_state(q).tobytes() == ref[q].tobytes(), q


def test_more_boxes_than_the_kernel_argument_table():
    """Beyond 64 boxes the tile and block kernels find their box through the
    device table instead of the kernel arguments; 70 boxes of five shapes,
    each equal to its shape stepped alone."""
    kw = MIXED_VARIANTS["ns_wall"]
    shapes = [(2, 2, 2, 2.0, 2.0, 2.0), (3, 2, 2, 3.0, 2.0, 2.0), (2, 3, 3, 2.0, 3.0, 3.0), (4, 2, 2, 4.0, 2.0, 2.0),
              (3, 3, 3, 3.0, 3.0, 3.0)]
    dts = [0.003, 0.0025, 0.0035, 0.002, 0.003]
    inits = [smooth_state(*x, seed=90 + q, amp=0.04, dtype=np.float32) for q, x in enumerate(shapes)]
    ref = []
    for q, x in enumerate(shapes):
        with od.Solver(od.params(*x, **kw)) as s:
            s.set_state(inits[q])
            s.step(dts[q], 2)
            ref.append(s.get_state())
    boxes = [shapes[q % 5] for q in range(70)]
    announce(test="seventy_boxes")
    with od.Batch(batch_params(kw), boxes) as b:
        for q in range(70):
            b.set_state(q, inits[q % 5])
        b.step([dts[q % 5] for q in range(70)], 2)
        for q in range(70):
            assert b.get_state(q).tobytes() == ref[q % 5].tobytes(), q

"""Batched runs (hitx_batch_*) against single-instance runs and the oracle.

Riemann instances in a batch must reproduce the single-instance solver bit
for bit (same kernels, same fixed dt, no host-side scalars).  Forced HIT
instances differ from the single-instance driver only through summation
order of the forcing statistics and the device exp/pow of the control
scalars (relative 1e-16 per step), so they are held to 1e-12 after a few
steps; against the oracle the sensitivity-based bound applies.
"""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def host(t):
    return t.detach().cpu().numpy()


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def test_riemann_batch_matches_single_instances():
    from hit_solver_amd.batch import BatchRunner
    from hit_solver_amd.riemann import RiemannConfig3, RiemannConfig3Offset08, RiemannConfig6, RiemannSolver

    cfgs = [RiemannConfig3(nx=64, ny=48, shear_threshold=None, progress_every=10**9),
            RiemannConfig6(nx=48, ny=64, shear_threshold=None, progress_every=10**9),
            RiemannConfig3Offset08(nx=56, ny=56, shear_threshold=None, progress_every=10**9, backend="hip")]
    b = BatchRunner(cfgs, max_steps=7)
    b.run()
    ctl = b.control()
    for i, c in enumerate(cfgs):
        s = RiemannSolver(c)
        s.run(max_steps=7)
        assert ctl[i, 4] == 7 and ctl[i, 3] == s.dt and ctl[i, 7] == s.hv_applications
        np.testing.assert_array_equal(host(b.q[i]), host(s.q))


def test_riemann_batch_split_sensor_bitexact():
    """The split sensor (undilated mask by column walk or on tiles, then the
    dilation) gives the same mask as the single plan: the batch reproduces
    single instances bit for bit (both reduce the stage on the same tiles)."""
    from hit_solver_amd.batch import BatchRunner
    from hit_solver_amd.riemann import RiemannConfig3, RiemannConfig6, RiemannSolver

    cfgs = [RiemannConfig3(nx=64, ny=48, shear_threshold=None, progress_every=10**9),
            RiemannConfig6(nx=40, ny=72, shear_threshold=None, progress_every=10**9)]
    for form in (2, 3):
        b = BatchRunner(cfgs, max_steps=5)
        b.tune(2, 8, 0, form)
        b.run()
        for i, c in enumerate(cfgs):
            s = RiemannSolver(c)
            s.run(max_steps=5)
            np.testing.assert_array_equal(host(b.q[i]), host(s.q))


def test_hit_batch_matches_single_instances():
    from hit_solver_amd.batch import BatchRunner
    from hit_solver_amd.hit2d import HIT2DConfig, HITSolver

    base = dict(cfl=0.05, initial_kmin=3, initial_kmax=5, forcing_kmin=3.0, forcing_kmax=5.0, viscosity=7.5e-4,
                forcing_correlation_time=0.5, mach_control=True, large_scale_drag=0.10, cooling_time=5.0,
                hyperviscosity_interval=2, tfinal=1.0, max_steps=5)
    cfgs = [HIT2DConfig(nx=32, ny=32, target_mach=0.1, forcing_seed=0, **base),
            HIT2DConfig(nx=48, ny=40, target_mach=0.25, forcing_seed=1, **base),
            HIT2DConfig(nx=64, ny=64, target_mach=0.4, forcing_seed=2, **base)]
    b = BatchRunner(cfgs)
    b.run()
    ctl = b.control()
    for i, c in enumerate(cfgs):
        s = HITSolver(c)
        s.run()
        assert ctl[i, 4] == s.steps
        assert ctl[i, 1] == pytest.approx(s.time, rel=1e-14)
        assert ctl[i, 12] == pytest.approx(s.info["alpha"], rel=1e-11)
        for k in range(4):
            assert rel_l2(host(b.q[i])[k], host(s.q)[k]) <= 1e-12, (i, k)


def test_hit_batch_mixed_hyperviscosity_schedules():
    """Instances with hyperviscosity every 2, 3 and 4 steps in one batch: on
    steps 1 and 5 no instance runs it (every post-step launch exits at the
    batch-wide word), on the others only some do."""
    from hit_solver_amd.batch import BatchRunner
    from hit_solver_amd.hit2d import HIT2DConfig, HITSolver

    base = dict(cfl=0.05, initial_kmin=3, initial_kmax=5, forcing_kmin=3.0, forcing_kmax=5.0, viscosity=7.5e-4,
                forcing_correlation_time=0.5, mach_control=True, large_scale_drag=0.10, cooling_time=5.0,
                tfinal=1.0, max_steps=6)
    cfgs = [HIT2DConfig(nx=32, ny=32, target_mach=0.1, forcing_seed=3, hyperviscosity_interval=2, **base),
            HIT2DConfig(nx=48, ny=40, target_mach=0.25, forcing_seed=4, hyperviscosity_interval=3, **base),
            HIT2DConfig(nx=40, ny=48, target_mach=0.4, forcing_seed=5, hyperviscosity_interval=4, **base)]
    print("[batch] mixed hyperviscosity schedules 2/3/4, 6 steps", flush=True)
    b = BatchRunner(cfgs)
    b.run()
    ctl = b.control()
    for i, (c, napp) in enumerate(zip(cfgs, (3, 2, 1))):
        s = HITSolver(c)
        s.run()
        assert ctl[i, 4] == s.steps == 6
        assert ctl[i, 7] == napp
        for k in range(4):
            assert rel_l2(host(b.q[i])[k], host(s.q)[k]) <= 1e-12, (i, k)


def test_hit_batch_vs_oracle():
    from hit_solver_amd.batch import BatchRunner
    from hit_solver_amd.hit2d import HIT2DConfig
    from oracle import hit as OH

    kw = dict(nx=48, ny=32, cfl=0.05, target_mach=0.25, initial_kmin=3, initial_kmax=5, forcing_kmin=3.0,
              forcing_kmax=5.0, viscosity=7.5e-4, forcing_correlation_time=0.5, mach_control=True,
              large_scale_drag=0.10, cooling_time=5.0, tfinal=1.0, max_steps=6, forcing_seed=7)
    b = BatchRunner([HIT2DConfig(**kw)])
    b.run()
    ref = OH.HITRun(OH.HITParams(**kw))
    ref.run()
    for k in range(4):
        assert rel_l2(host(b.q[0])[k], ref.q[k]) <= 1e-10, k


def assert_normals_match(got, ref):
    """Bit-exact, except values from the ziggurat tail (|x| > r = 3.654...),
    which go through log1p: the device's log1p and the host libm's may differ
    in the last bit (1 ulp)."""
    tail = np.abs(ref) > 3.6541528853610088
    np.testing.assert_array_equal(got[~tail], ref[~tail])
    np.testing.assert_array_max_ulp(got[tail], ref[tail], maxulp=1)


def test_device_normals_match_numpy():
    from hit_solver_amd.rng import DeviceNormalStream

    for seed, sizes in [(1235, [1, 7, 100, 4096]), (0, [262144, 262144]), (63, [37 * 41, 5000, 3])]:
        g = np.random.default_rng(seed)
        d = DeviceNormalStream(seed)
        for n in sizes:
            ref = g.standard_normal(n)
            assert_normals_match(host(d.normals(n)), ref)
        s = g.bit_generator.state["state"]
        assert d.host_state() == (s["state"], s["inc"])


def test_device_normals_serial_form_agrees():
    from hit_solver_amd.rng import DeviceNormalStream

    a, b = DeviceNormalStream(99), DeviceNormalStream(99)
    for n in (3000, 70000):
        np.testing.assert_array_equal(host(a.normals(n)), host(b.normals(n, serial=True)))
    assert a.host_state() == b.host_state()


def test_hit_batch_tuned_tiles_agree():
    """Tile heights change only the summation order of the column-walk
    reductions: a tuned batch agrees with the default one to rounding."""
    from hit_solver_amd.batch import BatchRunner
    from hit_solver_amd.hit2d import HIT2DConfig

    kw = dict(cfl=0.05, initial_kmin=3, initial_kmax=5, forcing_kmin=3.0, forcing_kmax=5.0, viscosity=7.5e-4,
              mach_control=True, large_scale_drag=0.10, cooling_time=5.0, tfinal=1.0, max_steps=4)
    cfgs = [HIT2DConfig(nx=64, ny=48, target_mach=0.25, forcing_seed=3, **kw),
            HIT2DConfig(nx=40, ny=72, target_mach=0.4, forcing_seed=4, **kw)]
    a, b = BatchRunner(cfgs), BatchRunner(cfgs)
    a.tune(2, 8, 1, 1)
    b.tune(8, 32, 2, 2)
    c = BatchRunner(cfgs)
    c.tune(4, 16, 2, 3)
    a.run()
    b.run()
    c.run()
    for i in range(2):
        for k in range(4):
            assert rel_l2(host(b.q[i])[k], host(a.q[i])[k]) <= 1e-12, (i, k)
            assert rel_l2(host(c.q[i])[k], host(a.q[i])[k]) <= 1e-12, (i, k)


def test_batch_graph_replay_matches_eager():
    """A captured batch step replayed k times equals k eager steps bit for bit
    (HIT with device noise and control, and Riemann), with eager work and
    synchronisations between the replays."""
    from hit_solver_amd.batch import BatchRunner
    from hit_solver_amd.hit2d import HIT2DConfig
    from hit_solver_amd.riemann import RiemannConfig3, RiemannConfig6

    kw = dict(cfl=0.05, initial_kmin=3, initial_kmax=5, forcing_kmin=3.0, forcing_kmax=5.0, viscosity=7.5e-4,
              mach_control=True, large_scale_drag=0.10, cooling_time=5.0, tfinal=1.0, hyperviscosity_interval=2)
    hit = [HIT2DConfig(nx=64, ny=64, target_mach=0.25, forcing_seed=5, **kw),
           HIT2DConfig(nx=48, ny=40, target_mach=0.1, forcing_seed=6, **kw)]
    rie = [RiemannConfig3(nx=64, ny=64, shear_threshold=None, progress_every=10**9),
           RiemannConfig6(nx=48, ny=56, shear_threshold=None, progress_every=10**9)]
    for cfgs in (hit, rie):
        a, b = BatchRunner(cfgs), BatchRunner(cfgs)
        b.capture(2)
        for _ in range(3):
            a.step()
            a.step()
            b.replay(1)
            torch.cuda.synchronize()
            _ = torch.ones(8, device="cuda").sum().item()  # eager work between replays
            for i in range(len(cfgs)):
                np.testing.assert_array_equal(host(b.q[i]), host(a.q[i]))
            np.testing.assert_array_equal(b.control(), a.control())


def test_hit_batch_walk_and_tile_drag_reductions_agree():
    """A batch above 4096 sensor tiles reduces the drag stage by column walk;
    a batch of one instance reduces it on tiles.  Only the summation order of
    the stage sums differs: the same instance agrees to rounding."""
    from hit_solver_amd.batch import BatchRunner
    from hit_solver_amd.hit2d import HIT2DConfig

    print("[batch] walk vs tile drag reductions, 33 x 256^2 FP64", flush=True)
    cfgs = [HIT2DConfig.production_512(nx=256, ny=256, forcing_seed=s, target_mach=0.25, tfinal=1e9, max_steps=3,
                                       dtype="fp64") for s in range(33)]
    big = BatchRunner(cfgs)  # 33 * 8 * 16 = 4224 tiles: column walk
    big.tune(16, 16, 2, 3)
    big.run()
    for i in (0, 32):
        one = BatchRunner([cfgs[i]])  # 128 tiles: tile form
        one.tune(16, 16, 2, 3)
        one.run()
        for k in range(4):
            assert rel_l2(host(big.q[i])[k], host(one.q[0])[k]) <= 1e-12, (i, k)


@pytest.mark.parametrize("dtype", ["fp32", "fp64"])
def test_slab_y_sweep_bitexact(dtype):
    """Line form 3 reads the y-lines from the row-major state through an LDS
    slab (enforced on load) instead of the transposed copy: the same
    arithmetic, so the batch is bit-identical to line form 2.  Grids cover
    y-lines of 1, 3, 4 and 7 wavefront chunks (L = ny/64) and column counts
    that leave the last slab block partly empty; one instance's x-lines are
    not a multiple of 64 (block form for x)."""
    from hit_solver_amd.batch import BatchRunner
    from hit_solver_amd.hit2d import HIT2DConfig
    from hit_solver_amd.riemann import RiemannConfig3, RiemannConfig6

    kw = dict(cfl=0.05, initial_kmin=3, initial_kmax=5, forcing_kmin=3.0, forcing_kmax=5.0, viscosity=7.5e-4,
              mach_control=True, large_scale_drag=0.10, cooling_time=5.0, tfinal=1.0, max_steps=3,
              hyperviscosity_interval=2, dtype=dtype)
    hit = [HIT2DConfig(nx=100, ny=64, target_mach=0.25, forcing_seed=1, **kw),
           HIT2DConfig(nx=192, ny=192, target_mach=0.1, forcing_seed=2, **kw),
           HIT2DConfig(nx=68, ny=256, target_mach=0.4, forcing_seed=3, **kw),
           HIT2DConfig(nx=448, ny=448, target_mach=0.25, forcing_seed=4, **kw)]
    rie = [RiemannConfig3(nx=100, ny=128, shear_threshold=None, progress_every=10**9, dtype=dtype),
           RiemannConfig6(nx=192, ny=320, shear_threshold=None, progress_every=10**9, dtype=dtype)]
    for cfgs in (hit, rie):
        print(f"[batch] slab y-sweep vs wave form, {len(cfgs)} instances {dtype}", flush=True)
        a, b = BatchRunner(cfgs, max_steps=3), BatchRunner(cfgs, max_steps=3)
        a.tune(4, 16, 2, 3)
        b.tune(4, 16, 3, 3)
        a.run()
        b.run()
        for i in range(len(cfgs)):
            np.testing.assert_array_equal(host(b.q[i]), host(a.q[i]), err_msg=f"instance {i}")
        np.testing.assert_array_equal(b.control(), a.control())


@pytest.mark.parametrize("dtype", ["fp64", "fp32"])
@pytest.mark.parametrize("stage_form,line_form", [(1, 2), (2, 2), (2, 1)])
def test_mixed_hit_riemann_batch_bitexact(dtype, stage_form, line_form):
    """Forced HIT with and without the drag source, with different sensor
    widths, and the three Riemann problems (hybrid and WENO-only schemes) in
    one batch: every instance is bit-identical to the same instance in a batch
    of its own kind.  The line sweeps launch once per periodicity and skip
    the other's instances; instances without drag reduce the stage on tiles
    whatever form the drag instances use."""
    from hit_solver_amd.batch import BatchRunner
    from hit_solver_amd.hit2d import HIT2DConfig
    from hit_solver_amd.riemann import RiemannConfig3, RiemannConfig3Offset08, RiemannConfig6

    kw = dict(cfl=0.05, initial_kmin=3, initial_kmax=5, forcing_kmin=3.0, forcing_kmax=5.0, viscosity=7.5e-4,
              mach_control=True, cooling_time=5.0, tfinal=1.0, hyperviscosity_interval=2, dtype=dtype)
    drag = [HIT2DConfig(nx=64, ny=64, target_mach=0.25, forcing_seed=1, large_scale_drag=0.10, **kw),
            HIT2DConfig(nx=128, ny=96, target_mach=0.4, forcing_seed=2, large_scale_drag=0.10, **kw)]
    nodrag = [HIT2DConfig(nx=48, ny=64, target_mach=0.1, forcing_seed=3, large_scale_drag=0.0, sensor_width=2, **kw)]
    rie = [RiemannConfig3(nx=64, ny=48, shear_threshold=None, progress_every=10**9, dtype=dtype),
           RiemannConfig6(nx=48, ny=64, shear_threshold=None, progress_every=10**9, dtype=dtype),
           RiemannConfig3Offset08(nx=56, ny=56, shear_threshold=None, progress_every=10**9, dtype=dtype,
                                  scheme="weno")]
    groups = [drag, nodrag, rie]
    order = [(0, 0), (2, 0), (1, 0), (2, 1), (0, 1), (2, 2)]  # (group, index) of each mixed-batch slot
    steps = 6

    def run(cfgs):
        b = BatchRunner(cfgs, max_steps=steps)
        b.tune(4, 16, line_form, 3)
        b.set_stage_form(stage_form)
        b.run()
        return b

    print(f"[batch] mixed HIT + Riemann {dtype}, stage form {stage_form}, line form {line_form}", flush=True)
    mixed = run([groups[g][i] for g, i in order])
    alone = [run(g) for g in groups]
    ctl = mixed.control()
    for slot, (g, i) in enumerate(order):
        np.testing.assert_array_equal(host(mixed.q[slot]), host(alone[g].q[i]), err_msg=f"slot {slot} = {g},{i}")
        np.testing.assert_array_equal(ctl[slot], alone[g].control()[i])
        assert ctl[slot, 4] == steps


def test_fused_x_sweep_update_bitexact():
    """Line form 4 runs each row panel's x-sweep and stage update in one
    launch (FP32, periodic grids, wave-form lines): the same arithmetic as the
    separate x-sweep and update of line form 2, so the batch is
    bit-identical, with hyperviscosity steps (the last stage writing S) and
    the other steps (writing Q in place) both covered."""
    from hit_solver_amd.batch import BatchRunner
    from hit_solver_amd.hit2d import HIT2DConfig

    kw = dict(cfl=0.05, initial_kmin=3, initial_kmax=5, forcing_kmin=3.0, forcing_kmax=5.0, viscosity=7.5e-4,
              mach_control=True, large_scale_drag=0.10, cooling_time=5.0, tfinal=1.0, max_steps=4,
              hyperviscosity_interval=2, dtype="fp32")
    cfgs = [HIT2DConfig(nx=64, ny=128, target_mach=0.25, forcing_seed=1, **kw),
            HIT2DConfig(nx=192, ny=192, target_mach=0.1, forcing_seed=2, **kw),
            HIT2DConfig(nx=448, ny=64, target_mach=0.4, forcing_seed=3, **kw),
            HIT2DConfig(nx=512, ny=256, target_mach=0.25, forcing_seed=4, **kw)]
    print(f"[batch] fused x-sweep + update vs wave form, {len(cfgs)} instances fp32", flush=True)
    a, b = BatchRunner(cfgs, max_steps=4), BatchRunner(cfgs, max_steps=4)
    a.tune(4, 16, 2, 3)
    b.tune(4, 16, 4, 3)
    assert not a.forms()["xfused"] and b.forms()["xfused"]
    a.run()
    b.run()
    for i in range(len(cfgs)):
        np.testing.assert_array_equal(host(b.q[i]), host(a.q[i]), err_msg=f"instance {i}")
    np.testing.assert_array_equal(b.control(), a.control())
    # a batch with a grid that is not periodic keeps the separate launches
    from hit_solver_amd.riemann import RiemannConfig3

    c = BatchRunner([cfgs[0], RiemannConfig3(nx=64, ny=64, shear_threshold=None, progress_every=10**9, dtype="fp32")])
    c.tune(4, 16, 4, 3)
    assert not c.forms()["xfused"]


def test_batch_serial_noise_draws_agree():
    """A batch whose noise draws all take the serial path (normals written to
    the noise buffer, projected by the projection kernel) agrees with the
    parallel draw, which projects the normals as it scatters them: the same
    normals, the projection summed in another order."""
    from hit_solver_amd.batch import BatchRunner
    from hit_solver_amd.hit2d import HIT2DConfig

    kw = dict(cfl=0.05, initial_kmin=3, initial_kmax=5, forcing_kmin=3.0, forcing_kmax=5.0, viscosity=7.5e-4,
              mach_control=True, large_scale_drag=0.10, cooling_time=5.0, tfinal=1.0, max_steps=3, dtype="fp64")
    cfgs = [HIT2DConfig(nx=64, ny=64, target_mach=0.25, forcing_seed=21, **kw),
            HIT2DConfig(nx=96, ny=80, target_mach=0.1, forcing_seed=22, **kw)]
    a, b = BatchRunner(cfgs, max_steps=3), BatchRunner(cfgs, max_steps=3)
    b.serial_draws(True)
    a.run()
    b.run()
    ca, cb = a.control(), b.control()
    assert (ca[:, 29] == 0.0).all() and (cb[:, 29] == 1.0).all()
    for i in range(len(cfgs)):
        for k in range(4):
            e = rel_l2(host(b.q[i])[k], host(a.q[i])[k])
            assert e <= 1e-12, (i, k, e)


def test_batch_noise_draws_stay_parallel():
    """The noise stream of forcing seed 61 at 512^2 (default_rng(62)) holds a
    ziggurat tail attempt that spills 8 words past a 32-word chunk in its
    292nd request, which chunk maps of 8 entry offsets sent to the serial
    path (about 60 ms); with 16 entry offsets every draw of the first 300
    steps stays on the parallel path."""
    from hit_solver_amd.batch import BatchRunner
    from hit_solver_amd.hit2d import HIT2DConfig

    cfg = HIT2DConfig.production_512(nx=512, ny=512, forcing_seed=61, target_mach=0.25, mach_control_target=0.25,
                                     dtype="fp32", tfinal=1e9)
    b = BatchRunner([cfg])
    print("[batch] 300 steps of seed 61 at 512^2, serial-path draws counted", flush=True)
    serial = []
    for s in range(300):
        b.step()
        if b.control()[0, 29] != 0.0:
            serial.append(s)
    assert serial == [], serial


def test_batch_time_updates_counts_launches():
    """The update timing entry point launches the three stage updates per
    repetition and periodicity present and returns a positive time per launch;
    a mixed batch (periodic HIT and non-periodic Riemann) launches both
    update kernels."""
    from hit_solver_amd.batch import BatchRunner
    from hit_solver_amd.hit2d import HIT2DConfig
    from hit_solver_amd.riemann import RiemannConfig3

    hit = HIT2DConfig(nx=64, ny=64, cfl=0.05, initial_kmin=3, initial_kmax=5, forcing_kmin=3.0, forcing_kmax=5.0,
                      viscosity=7.5e-4, tfinal=1.0)
    rie = RiemannConfig3(nx=64, ny=64, shear_threshold=None, progress_every=10**9)
    for cfgs, kinds in (([hit], 1), ([hit, rie], 2)):
        b = BatchRunner(cfgs)
        b.step()
        per_launch_ms, launches = b.time_updates(reps=2)
        assert launches == 3 * 2 * kinds
        assert np.isfinite(per_launch_ms) and 0.0 < per_launch_ms < 50.0, per_launch_ms
        with pytest.raises(Exception):
            b.time_updates(reps=0)

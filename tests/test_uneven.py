"""Heterogeneous batches: the host planner groups a mixed population into kernel classes
(no device needed), and on the GPU every drop of a graph-replayed mixed launch reproduces
the single-link launch of the same drop bit for bit."""

import ctypes

import numpy as np
import pytest


def test_population_covers_the_mix():
    from hermespy_amd.uneven import draw_population

    pop = draw_population(4000, seed=3)
    ants = {s.antennas for s in pop}
    ns = {s.num_subcarriers for s in pop}
    assert ants == {1, 2, 4, 8} and ns == {128, 256, 512, 1024, 2048}
    assert any(s.antennas == 8 and s.num_subcarriers == 2048 for s in pop)
    assert not any(s.antennas == 8 and s.num_subcarriers == 2048 for s in draw_population(500, 3, ((8, 2048),)))
    assert {s.modulation_order for s in pop} == {4, 16, 64}
    paths = np.array([s.num_paths for s in pop])
    assert paths.min() == 5 and paths.max() == 23
    dop = np.array([s.doppler_hz for s in pop])
    assert 10 <= dop.min() and dop.max() <= 900


def test_planner_groups_by_kernel_class(native_lib):
    """hpx_uneven_prepare on the host: classes are contiguous runs of one (antennas, threads, pairs,
    lean), heavier classes first, every drop placed once, LDS = max over the class; 8x8 drops at
    2048 subcarriers form lean classes within one CU's LDS."""

    from hermespy_amd import _lib
    from hermespy_amd.uneven import MAX_CLASSES, _c_link, draw_population

    pop = draw_population(300, seed=11)
    links = [s.link() for s in pop]
    n = len(links)
    c_links = (_lib.HpxLink * n)(*[_c_link(c) for c in links])
    tables = (_lib.HpxDropTables * n)()
    ws = [int(native_lib.hpx_uneven_workspace_elements(ctypes.byref(c_links[i]))) for i in range(n)]
    for i in range(n):
        tables[i].delays = 0x1000 + 256 * i  # host planning never dereferences device pointers
        tables[i].path_gains = 0x100000 + 512 * i
        tables[i].workspace = 0x10000000 + 0x1000000 * i if ws[i] else None
    idx = np.arange(n, dtype=np.uint32) + 17
    plan_bytes = int(native_lib.hpx_uneven_plan_bytes())
    plans = np.zeros(n * plan_bytes, dtype=np.uint8)
    order = np.empty(n, dtype=np.int64)
    classes = (_lib.HpxUnevenClass * MAX_CLASSES)()
    nc = _lib.c_int(0)
    _lib.check(native_lib.hpx_uneven_prepare(c_links, tables, idx.ctypes.data, n, ctypes.c_uint64(5), 0x2000, 0x3000,
                                             plans.ctypes.data, order.ctypes.data, classes, MAX_CLASSES,
                                             ctypes.byref(nc)))
    assert sorted(order.tolist()) == list(range(n))
    seen = lean_classes = 0
    for k in range(nc.value):
        c = classes[k]
        assert c.first == seen and c.count > 0
        members = order[c.first:c.first + c.count]
        shapes = {(links[i].num_tx, links[i].num_subcarriers) for i in members}
        assert len(shapes) == 1 and next(iter(shapes))[0] == c.num_tx
        lean = next(iter(shapes)) == (8, 2048)
        assert c.lean == int(lean) and all((ws[i] > 0) == lean for i in members)
        if lean:
            lean_classes += 1
            assert 0 < c.lds_bytes <= 160 * 1024
        else:
            # the single-link kernels of 512 threads keep the full table of realisation uniforms,
            # which the heterogeneous kernels replace by the 100 spatial ones
            lds = [int(native_lib.hpx_run_batch_lds_bytes(ctypes.byref(c_links[i]))) for i in members]
            if c.threads <= 256:
                assert c.lds_bytes == max(lds)
            else:
                assert 0 < c.lds_bytes <= max(lds)
        assert c.threads * c.pairs * 2 == links[members[0]].num_subcarriers
        seen += c.count
    assert seen == n and lean_classes >= 1


def test_lean_drops_need_a_workspace(native_lib):
    """An 8x8 link at 2048 subcarriers exceeds one CU's LDS: the single-link path refuses it, the
    heterogeneous planner sizes its workspace (Taylor table of every path and block, then N
    twiddles and the N - 1 entries of the FFT pass table) and refuses the drop without one."""

    from hermespy_amd import _lib
    from hermespy_amd.uneven import MAX_CLASSES, DropShape, _c_link

    big = _c_link(DropShape(8, 2048, 23, 2.0, 100.0, 16, 10.0).link())
    assert native_lib.hpx_run_batch_lds_bytes(ctypes.byref(big)) == 0
    ws = native_lib.hpx_uneven_workspace_elements(ctypes.byref(big))
    tables_len = 2048 + 2047
    assert ws > tables_len and (ws - tables_len) % (23 * 4) == 0  # 4 coefficients of the cubic
    fewer = _c_link(DropShape(8, 2048, 5, 2.0, 100.0, 16, 10.0).link())
    assert (native_lib.hpx_uneven_workspace_elements(ctypes.byref(fewer)) - tables_len) * 23 == (ws - tables_len) * 5
    for shape in ((8, 1024), (4, 2048), (1, 128)):
        small = _c_link(DropShape(shape[0], shape[1], 23, 2.0, 100.0, 16, 10.0).link())
        assert native_lib.hpx_uneven_workspace_elements(ctypes.byref(small)) == 0
    bad = _c_link(DropShape(3, 256, 5, 2.0, 100.0, 16, 10.0).link())
    assert native_lib.hpx_uneven_workspace_elements(ctypes.byref(bad)) == -1
    with pytest.raises(ValueError, match="not instantiated"):
        _lib.check_count(-1)
    c_links = (_lib.HpxLink * 1)(big)
    tables = (_lib.HpxDropTables * 1)()
    tables[0].delays, tables[0].path_gains = 0x1000, 0x2000
    idx = np.zeros(1, dtype=np.uint32)
    plans = np.zeros(int(native_lib.hpx_uneven_plan_bytes()), dtype=np.uint8)
    order = np.empty(1, dtype=np.int64)
    classes = (_lib.HpxUnevenClass * MAX_CLASSES)()
    nc = _lib.c_int(0)
    st = native_lib.hpx_uneven_prepare(c_links, tables, idx.ctypes.data, 1, ctypes.c_uint64(0), 0x10, 0x20,
                                       plans.ctypes.data, order.ctypes.data, classes, MAX_CLASSES, ctypes.byref(nc))
    with pytest.raises(ValueError, match="needs a workspace"):
        _lib.check(st)


def test_planner_rejects_bad_links(native_lib):
    from hermespy_amd import _lib
    from hermespy_amd.uneven import MAX_CLASSES, DropShape, _c_link

    bad = DropShape(3, 256, 5, 2.0, 100.0, 16, 10.0).link()  # 3x3 is not instantiated
    c_links = (_lib.HpxLink * 1)(_c_link(bad))
    tables = (_lib.HpxDropTables * 1)()
    tables[0].delays, tables[0].path_gains = 0x1000, 0x2000
    idx = np.zeros(1, dtype=np.uint32)
    plans = np.zeros(int(native_lib.hpx_uneven_plan_bytes()), dtype=np.uint8)
    order = np.empty(1, dtype=np.int64)
    classes = (_lib.HpxUnevenClass * MAX_CLASSES)()
    nc = _lib.c_int(0)
    st = native_lib.hpx_uneven_prepare(c_links, tables, idx.ctypes.data, 1, ctypes.c_uint64(0), 0x10, 0x20,
                                       plans.ctypes.data, order.ctypes.data, classes, MAX_CLASSES, ctypes.byref(nc))
    with pytest.raises(ValueError, match="not instantiated"):
        _lib.check(st)


def test_tuner_candidates():
    """Launch configurations the heterogeneous tuner tries: N, N/2, N/4 and N/8 threads within
    64..512, and the two FFT values per thread a class's kernels come in (default first)."""

    from hermespy_amd.uneven import fft_value_candidates, workgroup_candidates

    assert workgroup_candidates(128) == [128, 64]
    assert workgroup_candidates(512) == [512, 256, 128, 64]
    assert workgroup_candidates(1024) == [512, 256, 128]
    assert workgroup_candidates(2048) == [512, 256]
    assert fft_value_candidates(4, 512) == [12, 16] and fft_value_candidates(1, 512) == [12, 16]
    assert fft_value_candidates(8, 512) == [16, 8] and fft_value_candidates(8, 64) == [16, 8]
    assert fft_value_candidates(2, 256) == [16, 12] and fft_value_candidates(4, 128) == [16, 12]


def test_planner_splits_classes_by_fft_values(native_lib):
    """The FFT values per thread are part of a drop's kernel class: drops of one shape that ask for
    the default (0 or its value) and the alternative land in two classes that report their values;
    a value the class has no kernel for is refused."""

    from hermespy_amd import _lib
    from hermespy_amd.uneven import MAX_CLASSES, DropShape, _c_link, fft_value_candidates

    shapes = [DropShape(4, 2048, 9, 2.0, 100.0, 16, 10.0), DropShape(8, 1024, 7, 2.0, 300.0, 4, 20.0),
              DropShape(2, 256, 5, 1.0, 50.0, 64, 5.0)]

    def plan(requests):
        n = len(requests)
        c_links = (_lib.HpxLink * n)(*[_c_link(shapes[k].link(), 0, f) for k, f in requests])
        tables = (_lib.HpxDropTables * n)()
        for i in range(n):
            tables[i].delays, tables[i].path_gains = 0x1000 + 256 * i, 0x100000 + 512 * i
        idx = np.arange(n, dtype=np.uint32)
        plans = np.zeros(n * int(native_lib.hpx_uneven_plan_bytes()), dtype=np.uint8)
        order = np.empty(n, dtype=np.int64)
        classes = (_lib.HpxUnevenClass * MAX_CLASSES)()
        nc = _lib.c_int(0)
        st = native_lib.hpx_uneven_prepare(c_links, tables, idx.ctypes.data, n, ctypes.c_uint64(1), 0x2000, 0x3000,
                                           plans.ctypes.data, order.ctypes.data, classes, MAX_CLASSES,
                                           ctypes.byref(nc))
        return st, [classes[k] for k in range(nc.value)], order

    requests = []
    for k, shape in enumerate(shapes):
        default, alt = fft_value_candidates(shape.antennas, min(512, shape.num_subcarriers // 2))
        requests += [(k, 0), (k, alt), (k, default), (k, alt)]
    st, classes, order = plan(requests)
    _lib.check(st)
    assert len(classes) == 2 * len(shapes)
    for c in classes:
        members = order[c.first:c.first + c.count]
        ks = {requests[i][0] for i in members}
        assert len(ks) == 1 and c.count == 2
        k = ks.pop()
        default, alt = fft_value_candidates(shapes[k].antennas, c.threads)
        asked = {requests[i][1] or default for i in members}
        assert asked == {c.fft_values} and c.fft_values in (default, alt) and c.reserved == 0
    st, _, _ = plan([(0, 20)])
    with pytest.raises(ValueError, match="fft_values"):
        _lib.check(st)


@pytest.mark.gpu
def test_uneven_graph_replay_matches_single_link_launches(cuda):
    import torch

    from hermespy_amd.simulation import BatchedLink
    from hermespy_amd.uneven import UnevenBatch, draw_population

    pop = draw_population(96, seed=21)
    ub = UnevenBatch(pop, seed=99, first_drop=1000)
    ub.run()
    torch.cuda.synchronize()
    eager = ub.errors.cpu().numpy().copy()
    ub.capture()
    for rep in range(3):
        ub.errors.fill_(-1)
        torch.cuda.synchronize()
        ub.replay()
        torch.cuda.synchronize()
        assert np.array_equal(ub.errors.cpu().numpy(), eager), rep
    # each drop equals the plain single-link launch of the same drop
    for i in range(0, len(pop), 7):
        if (pop[i].antennas, pop[i].num_subcarriers) == (8, 2048):
            continue  # beyond the single-link path's LDS: checked against the oracle below
        cfg = pop[i].link()
        bl = BatchedLink(cfg)
        npw = torch.as_tensor([cfg.noise_power(pop[i].snr_db)], dtype=torch.float64, device=cuda)
        err = torch.empty(1, dtype=torch.int32, device=cuda)
        bl.run(99, 1000 + i, npw, err)
        torch.cuda.synchronize()
        assert int(err.item()) == int(eager[i]), (i, pop[i])
    assert 0 < eager.sum()
    # the classes forked over side streams give what one stream gives
    one = UnevenBatch(pop, seed=99, first_drop=1000, streams=1)
    one.run()
    torch.cuda.synchronize()
    assert np.array_equal(one.errors.cpu().numpy(), eager)


@pytest.mark.gpu
def test_uneven_autotuner_keeps_results(cuda):
    import torch

    from hermespy_amd.uneven import UnevenBatch, draw_population, tune

    pop = draw_population(64, seed=5)
    ub = UnevenBatch(pop, seed=1)
    ub.run()
    torch.cuda.synchronize()
    before = ub.errors.cpu().numpy().copy()
    info = tune(ub, reps=1)
    ub.run()
    torch.cuda.synchronize()
    assert np.array_equal(ub.errors.cpu().numpy(), before)
    assert info["choice"]


@pytest.mark.gpu
def test_alternative_fft_values_keep_results(cuda):
    """Every class on its alternative FFT values per thread (other row groups and per-pass register
    counts, the same arithmetic per element): error counts and equalised symbols equal the default
    kernels' bit for bit, lean 8x8 classes included."""

    import torch

    from hermespy_amd.uneven import DropShape, UnevenBatch, draw_population, fft_value_candidates

    pop = draw_population(80, seed=31) + [DropShape(8, 2048, 9, 2.0, 300.0, 16, 12.0)]
    print("alternative FFT values per thread vs default", flush=True)
    ub = UnevenBatch(pop, seed=4, first_drop=70, export_equalized=True)
    ub.run()
    torch.cuda.synchronize()
    errors = ub.errors.cpu().numpy().copy()
    eq = [t.cpu().numpy().copy() for t in ub.equalized]
    alt = {}
    for shape in pop:
        key = (shape.antennas, shape.num_subcarriers)
        threads = min(512, shape.num_subcarriers // 2)
        alt[key] = (threads, fft_value_candidates(shape.antennas, threads)[1])
    ub.prepare(alt)
    assert all(ub.classes[k].fft_values == alt[(ub.classes[k].num_tx, 2 * ub.classes[k].threads
                                                * ub.classes[k].pairs)][1] for k in range(ub.num_classes))
    ub.errors.fill_(-1)
    for t in ub.equalized:
        t.zero_()
    ub.run()
    torch.cuda.synchronize()
    assert np.array_equal(ub.errors.cpu().numpy(), errors)
    for i, t in enumerate(ub.equalized):
        assert np.array_equal(t.cpu().numpy(), eq[i]), i


@pytest.mark.gpu
def test_lean_drops_match_the_oracle(cuda):
    """8x8 drops at 2048 subcarriers (Taylor table and twiddles in device workspace, ideal CSI
    summed path by path) in a mixed batch: each drop's error count agrees with the float64
    oracle's up to symbols within rounding of a decision boundary, eagerly and on graph replay."""

    import torch

    from conftest import oracle_link
    from hermespy_amd.uneven import DropShape, UnevenBatch
    from oracle import drop as odrop

    lean = [DropShape(8, 2048, 23, 1.5, 400.0, 16, 12.0), DropShape(8, 2048, 5, 4.0, 60.0, 4, 6.0),
            DropShape(8, 2048, 14, 2.5, 850.0, 64, 24.0), DropShape(8, 2048, 9, 1.0, 200.0, 16, 30.0)]
    others = [DropShape(8, 1024, 11, 2.0, 300.0, 16, 15.0), DropShape(2, 2048, 7, 3.0, 100.0, 64, 21.0)]
    pop = lean + others
    print("lean 8x8 @ 2048 drops vs oracle", flush=True)
    ub = UnevenBatch(pop, seed=77, first_drop=500)
    assert ub.num_lean == len(lean)
    ub.run()
    torch.cuda.synchronize()
    eager = ub.errors.cpu().numpy().copy()
    ub.capture()
    for rep in range(2):
        ub.errors.fill_(-1)
        torch.cuda.synchronize()
        ub.replay()
        torch.cuda.synchronize()
        assert np.array_equal(ub.errors.cpu().numpy(), eager), rep
    for i, shape in enumerate(pop):
        cfg = shape.link()
        ref = odrop.run_drop(oracle_link(cfg), 77, 500 + i, shape.snr_db, faithful=False)
        assert ref.errors > 0
        assert abs(int(eager[i]) - ref.errors) <= max(2, int(2e-4 * ref.bits)), (i, shape, int(eager[i]), ref.errors)


def test_noalloc_planner_rejects_batches_beyond_its_scratch(native_lib):
    """Lean drops needing more than the scratch array's capacity are refused before any device work."""

    import ctypes

    from hermespy_amd import _lib
    from hermespy_amd.uneven import MAX_CLASSES, DropShape, _c_link

    cap = int(native_lib.hpx_uneven_noalloc_capacity())
    cfg = DropShape(8, 2048, 23, 1.5, 400.0, 16, 12.0).link()
    per = int(native_lib.hpx_uneven_workspace_elements(ctypes.byref(_c_link(cfg))))
    assert per > 0
    n = cap // per + 1
    c_links = (_lib.HpxLink * n)(*[_c_link(cfg)] * n)
    tables = (_lib.HpxDropTables * n)()
    for i in range(n):
        tables[i].delays, tables[i].path_gains = 0x10, 0x20
    idx = np.arange(n, dtype=np.uint32)
    plans = np.empty(n * int(native_lib.hpx_uneven_plan_bytes()), dtype=np.uint8)
    order = np.empty(n, dtype=np.int64)
    classes = (_lib.HpxUnevenClass * MAX_CLASSES)()
    nc = _lib.c_int(0)
    st = native_lib.hpx_uneven_prepare_noalloc(c_links, tables, idx.ctypes.data, n, ctypes.c_uint64(0), 0x30, 0x40,
                                               plans.ctypes.data, order.ctypes.data, classes, MAX_CLASSES,
                                               ctypes.byref(nc))
    with pytest.raises(ValueError, match="scratch array holds"):
        _lib.check(st)


@pytest.mark.gpu
def test_every_kernel_class_matches_the_oracle(cuda):
    """One drop of every (antennas, subcarriers) kernel class of the heterogeneous batch --
    1x1 .. 8x8 at 128 .. 2048 subcarriers, the 8x8 / 2048 drop in a lean class -- with QPSK,
    16-QAM and 64-QAM, against the float64 oracle: equalised symbols within 1e-4, exact demapping
    of the kernel's own symbols, error counts within the decision-boundary slack."""

    import torch

    from conftest import oracle_link
    from oracle_pool import oracle_links
    from parity import check_drop

    from hermespy_amd.uneven import DropShape, UnevenBatch

    shapes = []
    k = 0
    for a in (1, 2, 4, 8):
        for n in (128, 256, 512, 1024, 2048):
            shapes.append(DropShape(a, n, 5 + (7 * k) % 19, 1.0 + (k % 6), 10.0 + 45.0 * k, (4, 16, 64)[k % 3],
                                    float(3 * (k % 11))))
            k += 1
    print(f"{len(shapes)} kernel classes vs the oracle", flush=True)
    ub = UnevenBatch(shapes, seed=31, first_drop=4000, export_equalized=True)
    assert ub.num_lean == 1
    ub.run()
    torch.cuda.synchronize()
    err = ub.errors.cpu().numpy()
    eqs = [e.cpu().numpy() for e in ub.equalized]
    links = [oracle_link(s.link()) for s in shapes]
    ora = oracle_links(links, 31, [4000 + i for i in range(len(shapes))], [s.snr_db for s in shapes])
    for i, s in enumerate(shapes):
        check_drop(eqs[i], err[i], ora[i], s.modulation_order, 1e-4, (s.antennas, s.num_subcarriers))


def test_launch_schedule_covers_every_slot_once_and_balances():
    """schedule_launches cuts classes costlier than half a stream's share into drop ranges and
    assigns the pieces longest first to the least-loaded stream: every plan slot is launched
    exactly once, with its class's descriptor, and no stream exceeds the lightest by more than the
    largest piece (the LPT bound)."""

    from hermespy_amd import _lib
    from hermespy_amd.uneven import schedule_launches

    rng = np.random.default_rng(5)
    counts = rng.integers(1, 500, 20)
    firsts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    classes = [_lib.HpxUnevenClass(int(f), int(c), int(rng.choice([1, 2, 4, 8])), 256, 1, 0, 4096,
                                   int(rng.choice([8, 12, 16])), 0)
               for f, c in zip(firsts, counts)]
    cost = counts * rng.uniform(0.5, 20.0, counts.size)
    lanes, load = schedule_launches(classes, len(classes), cost, 4)
    slots = np.zeros(int(counts.sum()), dtype=np.int64)
    biggest = 0.0
    for lane in lanes:
        for pc in lane:
            k = int(np.searchsorted(firsts, pc.first, side="right") - 1)
            assert pc.first + pc.count <= firsts[k] + counts[k]
            assert (pc.num_tx, pc.threads, pc.lds_bytes) == (classes[k].num_tx, 256, 4096)
            assert (pc.pairs, pc.lean, pc.fft_values) == (classes[k].pairs, classes[k].lean, classes[k].fft_values)
            slots[pc.first:pc.first + pc.count] += 1
            biggest = max(biggest, cost[k] * pc.count / counts[k])
    assert np.all(slots == 1)
    assert abs(sum(load) - cost.sum()) < 1e-6 * cost.sum()
    assert max(load) - min(load) <= biggest + 1e-9
    assert biggest <= cost.sum() / (2 * 4) + cost.max() / counts.min()  # split pieces stay near the share
    # one stream still launches every slot once
    lanes1, _ = schedule_launches(classes, len(classes), cost, 1)
    assert sum(pc.count for pc in lanes1[0]) == counts.sum()

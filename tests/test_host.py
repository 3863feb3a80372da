"""CPU tests of the boundary and host logic: the C ABI loads and exports what
include/eplace.h declares, configuration defaults, the generator, the host
legalizer against the oracle's brute force, and loud failure without a GPU."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import _oracle as O
import eisenjoch_amd as ep
from eisenjoch_amd import _lib as L

ROOT = Path(__file__).resolve().parent.parent


def header_symbols():
    text = (ROOT / "include" / "eplace.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ep_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    syms = header_symbols()
    assert len(syms) >= 29
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(L.SIGNATURES), set(syms) ^ set(L.SIGNATURES)
    assert lib.ep_version() == 1


def test_struct_layouts_match_header(tmp_path):
    # the ctypes mirror must agree with the C compiler on every offset
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    structs = {"ep_design": L.EpDesign, "ep_config": L.EpConfig, "ep_iter_rec": L.EpIterRec,
               "ep_result": L.EpResult, "ep_serve_config": L.EpServeConfig,
               "ep_gen_params": L.EpGenParams, "ep_gen_design": L.EpGenDesign,
               "ep_phase_time": L.EpPhaseTime, "ep_serve_stats": L.EpServeStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "eplace.h"', "int main(void){"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f in py._fields_:
            cf = "lambda" if f[0] == "lambda_" else f[0]
            lines.append(f'printf("{cname}.{f[0]} %zu\\n", offsetof({cname}, {cf}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True)
    out = dict(l.rsplit(" ", 1) for l in subprocess.check_output([str(exe)], text=True).split("\n") if l)
    for cname, py in structs.items():
        assert int(out[cname]) == C.sizeof(py), cname
        for f in py._fields_:
            assert int(out[f"{cname}.{f[0]}"]) == getattr(py, f[0]).offset, (cname, f[0])


def test_config_defaults_are_the_reference():
    # config.rs:27-38 and mod.rs:31-43 (electro_place_tests.rs:97-109)
    c = L.EpConfig()
    L.lib().ep_config_default(C.byref(c))
    assert c.seed == 1 and c.wl_coeff == 0.5 and c.target_util == 0.7
    assert c.timing_driven == 0 and c.target_density == 1.0 and c.timing_weight == 0.0
    assert c.nesterov_step_size == 0.1 and c.max_iters == 500 and c.legalize_interval == 5
    assert c.min_iters == 20 and c.overflow_target == 0.1 and c.density_eta == 0.1
    assert c.grow_mult == 1.025 and c.grow_threshold == 50.0 and c.grow_add == 1.0
    assert c.field_mode == L.EP_FIELD_REFERENCE
    py = ep.ElectroPlaceCfg()
    assert (py.seed, py.wl_coeff, py.max_iters, py.legalize_interval) == (1, 0.5, 500, 5)
    assert py.to_c().grow_mult == 1.025


def test_context_creation_fails_loudly_without_gpu():
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a device is present")
    h = C.c_void_p()
    rc = L.lib().ep_create(C.byref(h), 0, 0)
    assert rc == L.EP_ERR_DEVICE
    with pytest.raises(ep.DeviceError):
        ep.Device()


def test_generator_config1_shape():
    d = ep.config1()
    assert (d.n_mov, d.n_fixed, d.n_nets, d.W, d.H, d.Bx, d.By) == (1500, 64, 1700, 64, 64, 64, 64)
    assert (d.cell_type == 1).sum() == 300
    deg = np.diff(d.net_off.astype(np.int64))
    assert deg.min() >= 2 and deg.max() <= 64
    frac2 = (deg == 2).mean()
    assert 0.5 < frac2 < 0.7
    # IO on the perimeter, integer tiles
    assert np.all((d.fixed_x == 0) | (d.fixed_x == 63) | (d.fixed_y == 0) | (d.fixed_y == 63))
    assert np.all(d.pin_cell < d.n_mov + d.n_fixed)
    assert ((d.pin_cell >= d.n_mov).sum()) == 64
    # initial positions: each cell on its own BEL of its type
    assert len(set(d.init_bel.tolist())) == d.n_mov
    assert np.all(d.bel_type[d.init_bel] == d.cell_type)
    np.testing.assert_array_equal(d.x0, d.bel_x[d.init_bel])
    # deterministic
    d2 = ep.config1()
    np.testing.assert_array_equal(d.pin_cell, d2.pin_cell)
    np.testing.assert_array_equal(d.x0, d2.x0)


def test_generator_initial_placement_is_the_reference_shuffle():
    # common.rs:79-113 with rng.rs:52-61: per type, shuffle the bucket's
    # BELs (BEL index order) with xorshift64 from the seed (0 -> 1), bind in order
    for seed in (0, 1, 42):
        d = ep.generate(300, 330, 20, n_fixed=8, n_types=2, type1_frac=0.25, seed=seed)
        st = C.c_uint64(max(seed, 1))
        for t in (0, 1):
            avail = np.nonzero(d.bel_type == t)[0].astype(np.uint32)
            O.lib().orc_rng_shuffle_u32(C.byref(st), O.P(avail), avail.size)
            cells = np.nonzero(d.cell_type == t)[0]
            np.testing.assert_array_equal(d.init_bel[cells], avail[: cells.size])


def test_generator_area_mix_and_high_fanout():
    d = ep.generate(20_000, 22_000, 170, n_fixed=100, Bx=128, By=128, n_clusters=100,
                    area_mix=True, n_high_fanout=5, hf_min=500, hf_max=600, seed=3)
    a = d.area
    assert set(np.unique(a).tolist()) == {1.0, 2.0, 4.0}
    assert 0.82 < (a == 1).mean() < 0.88
    deg = np.diff(d.net_off.astype(np.int64))
    assert (deg >= 500).sum() == 5


def test_batch_rule():
    assert ep.bins_for_side(53) == 64 and ep.bins_for_side(169) == 128
    assert ep.bins_for_side(378) == 512 and ep.bins_for_side(3000) == 512
    s = ep.batch_sizes(64)
    assert len(s) == 64 and min(s) >= 2000 and max(s) <= 100_000


@pytest.mark.parametrize("seed", range(6))
def test_host_legalizer_matches_reference_scan(seed):
    # ep_legalize (ring search) equals the reference's full bucket scan
    # (mod.rs:216-237) bit for bit, ties included
    rng = np.random.default_rng(seed)
    W = H = int(rng.integers(5, 14))
    T = 2
    bx = np.repeat(np.tile(np.arange(W), H), T).astype(np.int32)
    by = np.repeat(np.repeat(np.arange(H), W), T).astype(np.int32)
    bb = np.tile(np.arange(T), W * H).astype(np.int32)
    n = int(rng.integers(1, W * H))
    ct = rng.integers(0, T, n).astype(np.int32)
    if seed % 2:
        x = rng.integers(-2, 2 * W + 2, n) / 2.0  # half-integers: many exact ties
        y = rng.integers(-2, 2 * H + 2, n) / 2.0
    else:
        x = rng.uniform(-1, W, n)
        y = rng.uniform(-1, H, n)
    taken0 = (rng.uniform(size=bx.size) < 0.1).astype(np.uint8)
    rc, ref, dref = O.legalize(x, y, ct, bx, by, bb, taken0)
    if rc != 0:
        with pytest.raises(ep.NoBelsAvailable):
            ep.legalize(x, y, ct, bx, by, bb, taken0)
        return
    got, dgot = ep.legalize(x, y, ct, bx, by, bb, taken0)
    np.testing.assert_array_equal(got, ref)
    assert dgot == dref


def test_total_hpwl_matches_oracle():
    d = ep.config1()
    ix = d.bel_x[d.init_bel]
    iy = d.bel_y[d.init_bel]
    a = ep.total_hpwl_int(d, ix, iy)
    b = O.lib().orc_total_hpwl_int(O.Design(d).ref, O.P(ix.astype(np.int32)), O.P(iy.astype(np.int32)))
    assert a == b
    assert a == pytest.approx(O.hpwl(d, d.x0, d.y0))


def test_unknown_placer_is_rejected():
    ctx = ep.Context(ep.config1())
    with pytest.raises(ValueError):
        ctx.place(placer="annealing")


def test_design_bytes_sizes_an_arena():
    d = ep.config1()
    b64 = d.nbytes("fp64")
    b32 = d.nbytes("fast")
    assert b64 > b32 > 0


def test_bench_refuses_more_than_one_gpu():
    # the placer runs on one GPU; bench.py says so instead of measuring
    import subprocess
    import sys as _sys
    r = subprocess.run([_sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stdout


def test_bench_roofline_inputs():
    # algorithmic work per design follows the stated per-unit figures, and
    # the traffic of a group adds its kernels' calibrated counter bytes
    import json
    import sys as _sys
    _sys.path.insert(0, str(ROOT))
    import bench
    d = ep.sized_design(3000, 1)
    w = bench.algorithmic_work([d])
    N, F, E, P, B2 = d.n_mov, d.n_fixed, d.n_nets, d.n_pins, d.Bx * d.By
    assert w["wa"] == 4 * (E + 1) + 4 * P + 16 * N + 8 * F
    assert w["scatter"] == 12 * N + 4 * B2
    assert w["gather+update"] == 60 * N + 8 * B2
    assert w["dct_flops"] == 36 * d.Bx ** 3  # square grid: 18 Bx By (Bx + By)
    t = {"fetch_factor": 2.0, "write_factor": 1.0,
         "kernels": {"void ep::k_wa<float, true, false>(ep::BatchDev, int)": {"fetch_kib": 10.0, "write_kib": 4.0},
                     "void ep::k_wl_cell<float>(ep::BatchDev)": {"fetch_kib": 1.0, "write_kib": 1.0}}}
    import tempfile
    with tempfile.NamedTemporaryFile("w", suffix=".json", delete=False) as f:
        json.dump(t, f)
    assert bench.group_traffic(f.name, bench.KERNEL_GROUPS["wa"]) == 1024 * (24 + 3)
    assert bench.group_traffic(f.name, bench.KERNEL_GROUPS["gather+update"]) is None
    assert bench.group_traffic("/nonexistent.json", bench.KERNEL_GROUPS["wa"]) is None


def test_bench_kernel_groups_match_the_committed_profile():
    # every kernel pattern of a roofline group names a kernel of the committed
    # counter summary (a renamed or re-templated kernel would otherwise drop
    # the group's traffic to null without notice)
    import json
    import sys as _sys
    _sys.path.insert(0, str(ROOT))
    import bench
    t = json.loads((ROOT / "profiles" / "pmc_traffic_batched64.json").read_text())
    names = list(t["kernels"])
    for g in bench.KERNEL_GROUPS.values():
        for prefix in g["kernels"]:
            assert any(prefix in n for n in names), prefix
        assert bench.group_traffic(str(ROOT / "profiles" / "pmc_traffic_batched64.json"), g) > 0

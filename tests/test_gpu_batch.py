"""Batched independent systems (gse_batch_*) against the FP64 oracle run
system by system.

  * FP64: six uneven systems (n = 6, 7 and 8 supports, an anisotropic box
    and an empty system): energies, forces and potentials within 1e-10;
  * FP32: 32 water boxes of the batched workload (600..8000 charges,
    M in {24, 32, 40}): per-system force Erms <= 1e-5, energy within 1e-5;
  * a dense cluster that overflows its bin inside the batch;
  * the HIP-graph replay equals the eager step bit for bit, on each of
    several replays with other GPU work between them, and follows new
    particle data written in place;
  * the autotuner keeps a candidate shape and results stay correct.
"""
import numpy as np
import pytest
import torch

from gse_amd import api, systems
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

F64, F32 = torch.float64, torch.float32
SIG = 2.12


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    s = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / s) if s > 0 else float(np.max(np.abs(a - b)))


def small_systems():
    out = []
    out.append((systems.water_box(60, 16.0, seed=31), (16, 16, 16), (16.0,) * 3))  # n = 6
    out.append((systems.water_box(80, 17.0, seed=32), (20, 20, 20), (17.0,) * 3))  # n = 7
    out.append((systems.water_box(70, 18.0, seed=33), (24, 24, 24), (18.0,) * 3))  # n = 8
    out.append((np.zeros((0, 4)), (16, 16, 16), (16.0,) * 3))                        # empty
    aniso = systems.water_box(50, 18.0, seed=34) * np.array([1.0, 0.9, 1.2, 1.0])
    out.append((aniso, (18, 16, 24), (18.0, 16.2, 21.6)))
    out.append((systems.water_box(60, 16.0, seed=35), (16, 16, 16), (16.0,) * 3))  # same grid as #0
    return out


def run_oracle(sysl, rounded=None):
    refs = []
    for k, (x, dims, box) in enumerate(sysl):
        xx = x if rounded is None else rounded[k]
        refs.append(orc.gse_step(orc.Params(dims, box, SIG), xx, mode=0))
    return refs


def concat(sysl):
    return np.concatenate([x for x, _, _ in sysl], axis=0)


def check_against(sysl, refs, o, tol_f, tol_e, erms=False):
    f = o["force"].double().cpu().numpy().T
    e = o["energy"].cpu().numpy()
    pot = o["pot"].double().cpu().numpy() if "pot" in o else None
    p0 = 0
    for k, (x, _, _) in enumerate(sysl):
        n = len(x)
        r = refs[k]
        if n == 0:
            assert e[k] == 0.0
            continue
        fk = f[p0:p0 + n]
        if erms:
            assert orc.erms(r.force, fk) <= tol_f, k
        else:
            assert rel(fk, r.force) <= tol_f, k
            if pot is not None:
                assert rel(pot[p0:p0 + n], r.pot) <= tol_f, k
        assert abs(e[k] - r.energy) / abs(r.energy) <= tol_e, (k, e[k], r.energy)
        p0 += n


def make_batch(dt, sysl):
    return api.GSEBatch(dt, dt, [(dims, box, len(x)) for x, dims, box in sysl], SIG)


def test_batch_fp64_vs_oracle():
    sysl = small_systems()
    refs = run_oracle(sysl)
    b = make_batch(F64, sysl)
    x = torch.from_numpy(concat(sysl)).cuda()
    o = b.step(x, force=True, potential=True)
    torch.cuda.synchronize()
    check_against(sysl, refs, o, 1e-10, 1e-10)


def test_batch_fp32_water_boxes():
    bs = systems.batched_systems(32, seed=4)
    sysl = [(x, (M, M, M), (L, L, L)) for x, L, M in bs]
    xc = torch.from_numpy(concat(sysl)).to("cuda", F32)
    xr = xc.double().cpu().numpy()
    rounded, p0 = [], 0
    for x, _, _ in sysl:
        rounded.append(xr[p0:p0 + len(x)])
        p0 += len(x)
    refs = run_oracle(sysl, rounded)
    b = make_batch(F32, sysl)
    o = b.step(xc, force=True)
    torch.cuda.synchronize()
    check_against(sysl, refs, o, 1e-5, 1e-5, erms=True)


def test_batch_overflow_cluster_fp64():
    sysl = small_systems()[:3]
    rng = np.random.default_rng(41)
    cl = np.empty((600, 4))
    cl[:, :3] = 5.0 + rng.random((600, 3)) * 0.8
    cl[:, 3] = 0.25 + 0.5 * rng.random(600)
    sysl.insert(1, (cl, (16, 16, 16), (16.0,) * 3))
    refs = run_oracle(sysl)
    b = make_batch(F64, sysl)
    o = b.step(torch.from_numpy(concat(sysl)).cuda(), force=True, potential=True)
    torch.cuda.synchronize()
    check_against(sysl, refs, o, 1e-10, 1e-10)


def test_batch_graph_replays_match_eager():
    bs = systems.batched_systems(16, seed=7)
    sysl = [(x, (M, M, M), (L, L, L)) for x, L, M in bs]
    x = torch.from_numpy(concat(sysl)).to("cuda", F32)
    b = make_batch(F32, sysl)
    eager = b.step(x, force=True, potential=True)
    ef, ep, ee = (eager[k].clone() for k in ("force", "pot", "energy"))
    g = {}
    for rep in range(4):
        b.step(x, force=True, potential=True, graph=True, out=g)
        torch.cuda.synchronize()
        assert torch.equal(g["force"], ef), rep
        assert torch.equal(g["pot"], ep), rep
        assert torch.equal(g["energy"], ee), rep
        # unrelated eager work and an eager batched step between replays
        y = torch.randn(1 << 16, device="cuda").cumsum(0)
        b.step(x, force=True, potential=True)
        torch.cuda.synchronize()
        del y
    # the graph reads the particle buffer at replay time
    x2 = x.clone()
    x2[:, :3] += 0.37
    eager2 = b.step(x2, force=True)
    f2 = eager2["force"].clone()
    x.copy_(x2)
    b.step(x, force=True, potential=True, graph=True, out=g)
    torch.cuda.synchronize()
    assert torch.equal(g["force"], f2)


def test_batch_autotune_keeps_results():
    bs = systems.batched_systems(12, seed=9)
    sysl = [(x, (M, M, M), (L, L, L)) for x, L, M in bs]
    xc = torch.from_numpy(concat(sysl)).to("cuda", F32)
    xr = xc.double().cpu().numpy()
    rounded, p0 = [], 0
    for x, _, _ in sysl:
        rounded.append(xr[p0:p0 + len(x)])
        p0 += len(x)
    refs = run_oracle(sysl, rounded)
    b = make_batch(F32, sysl)
    cfg, ms = b.autotune(xc, reps=2)
    print("autotuned", cfg, f"{ms:.3f} ms/step", flush=True)
    assert cfg["bin_edge"] in (6, 8, 10)
    assert cfg["spread_block"] in (256, 512) and cfg["gather_block"] in (256, 512)
    assert ms > 0
    o = b.step(xc, force=True, graph=True)
    torch.cuda.synchronize()
    check_against(sysl, refs, o, 1e-5, 1e-5, erms=True)


def test_batch_bad_arguments():
    with pytest.raises(api._lib.GSEError):
        api.GSEBatch(F32, F64, [((16, 16, 16), 16.0, 10)], SIG)
    b = api.GSEBatch(F32, F32, [((16, 16, 16), 16.0, 10)], SIG)
    with pytest.raises(ValueError):
        b.step(torch.zeros((11, 4), device="cuda"))
    with pytest.raises(api._lib.GSEError):
        b.set_config(bin_edge=0)


def test_batch_fp64_brick_path_with_overflow():
    """Grids of whole 8-cell bins (>= 3 per axis) take the brick-store flush
    and tiled reduce; one system also overfills a bin, so the reduce adds the
    overflow particles' int64 grid.  FP64 within 1e-10 of the oracle."""
    rng = np.random.default_rng(43)
    cl = np.empty((900, 4))
    cl[:, :3] = 7.0 + rng.random((900, 3)) * 0.8
    cl[:, 3] = 0.25 + 0.5 * rng.random(900)
    cl = np.concatenate([cl, systems.water_box(100, 24.0, seed=44)])
    sysl = [(systems.water_box(200, 24.0, seed=45), (24, 24, 24), (24.0,) * 3),
            (cl, (24, 24, 24), (24.0,) * 3),
            (systems.water_box(300, 30.0, seed=46), (32, 32, 32), (30.0,) * 3),
            (systems.water_box(250, 26.0, seed=47), (32, 32, 32), (26.0,) * 3)]
    refs = run_oracle(sysl)
    b = make_batch(F64, sysl)
    o = b.step(torch.from_numpy(concat(sysl)).cuda(), force=True, potential=True)
    torch.cuda.synchronize()
    check_against(sysl, refs, o, 1e-10, 1e-10)


def _rounded_f32(sysl):
    xc = torch.from_numpy(concat(sysl)).to("cuda", F32)
    xr = xc.double().cpu().numpy()
    rounded, p0 = [], 0
    for x, _, _ in sysl:
        rounded.append(xr[p0:p0 + len(x)])
        p0 += len(x)
    return xc, rounded


def test_batch_every_candidate_shape():
    """Each launch shape the autotuner may pick (bin edge 6, 8, 10 x spread
    block 256, 512 x gather block 256, 512) gives FP32 results within Erms
    1e-5 of the oracle.  Edges 6 and 10 leave the brick store for the
    atomic flush, so both conversion paths of the batch are covered."""
    bs = systems.batched_systems(12, seed=9)
    sysl = [(x, (M, M, M), (L, L, L)) for x, L, M in bs]
    xc, rounded = _rounded_f32(sysl)
    refs = run_oracle(sysl, rounded)
    b = make_batch(F32, sysl)
    for edge in (6, 8, 10):
        for sb in (256, 512):
            for gb in (256, 512):
                print(f"shape edge={edge} spread={sb} gather={gb}", flush=True)
                b.set_config(bin_edge=edge, spread_block=sb, gather_block=gb)
                assert b.get_config() == dict(bin_edge=edge, spread_block=sb, gather_block=gb)
                o = b.step(xc, force=True, graph=(sb == gb))
                torch.cuda.synchronize()
                check_against(sysl, refs, o, 1e-5, 1e-5, erms=True)


NSAMPLE_PER_SYS = 64


def test_batch_256_water_boxes_graph_vs_oracle():
    """The batched workload at full size: 256 water boxes of 600..8000
    charges (1,130,376 in all), M in {24, 32, 40}, FP32, autotuned, one
    HIP-graph replay per step.  Every system is checked against the FP64
    oracle on the same float32-rounded particles: its energy (|dE|/E <=
    1e-5) and its reference forces on a fixed sample of 64 charges (Erms
    <= 1e-5, cpu_gse.cpp:131-146).  The oracle's spread, FFT solve and
    energy run on every system; its gather only on the sample."""
    bs = systems.batched_systems(256, seed=4)
    sysl = [(x, (M, M, M), (L, L, L)) for x, L, M in bs]
    xc, rounded = _rounded_f32(sysl)
    b = make_batch(F32, sysl)
    cfg, ms = b.autotune(xc, reps=2)
    print("autotuned", cfg, f"{ms:.3f} ms/step", flush=True)
    g = {}
    for _ in range(3):
        b.step(xc, force=True, graph=True, out=g)
    torch.cuda.synchronize()
    f = g["force"].double().cpu().numpy().T
    e = g["energy"].cpu().numpy()
    rng = np.random.default_rng(256)
    p0, worst_f, worst_e = 0, 0.0, 0.0
    for k, (x, dims, box) in enumerate(sysl):
        n = len(x)
        p = orc.Params(dims, box, SIG)
        xr = rounded[k]
        rho = orc.spread(p.dims, p.sigma1, p.rcut1, xr, p.box)
        phi = orc.fft_solve(0, p.dims, p.sigma2, p.box, rho)
        E = orc.energy(p.dims, p.ccelec, p.box, rho, phi)
        idx = np.sort(rng.choice(n, min(n, NSAMPLE_PER_SYS), replace=False))
        fs = orc.interpolate(p.dims, p.sigma1, p.rcut1, xr[idx], p.box, phi) * p.ccelec
        ef = orc.erms(fs, f[p0 + idx])
        ee = abs(e[k] - E) / abs(E)
        worst_f, worst_e = max(worst_f, ef), max(worst_e, ee)
        assert ef <= 1e-5, (k, ef)
        assert ee <= 1e-5, (k, ee)
        p0 += n
    print(f"256 systems: worst Erms {worst_f:.2e}, worst |dE|/E {worst_e:.2e}", flush=True)

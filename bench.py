"""Benchmark: FP32 clover Dslash (A^-1 D psi, one checkerboard) on a 48^3x96
lattice, reported as algorithmic HBM GB/s, plus the solver comparison of the
headline metric (mixed-precision clover solve to 1e-10 vs a pure-FP64
BiCGStab on the same GPU).

One step = one single-parity clover Dslash over all 5,308,416 target sites
with the inputs resident in HBM. Algorithmic bytes per target site: 8 links
(12 reals with two-row compression), one source spinor, one output spinor and
one 72-real clover block: (8*12 + 24 + 24 + 72) * 4 B = 864 B (FP32).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-solve]
"""
import argparse
import ctypes
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

LATT = [48, 48, 48, 96]
KAPPA = 0.1245
MASS = 1.0 / (2 * KAPPA) - 4.0   # alpha = 4 + m = 1/(2 kappa)
CSW = 1.0
HBM_PEAK_GBS = 8000.0
BYTES_SITE = {("f32", True): 864, ("f32", False): 1056, ("f64", False): 2112, ("f16", True): 448,
              ("bf16", True): 448}
FLOPS_SITE_CLOVER = 1824
RHS_TILE = []  # forced multi-RHS tiling (--rhs-tile), empty: autotuned


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def traffic_from_profiles():
    """HBM bytes per launch of the clover Dslash kernel from the committed
    rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated)."""
    path = os.path.join(ROOT, "profiles", "clover_dslash_f32c12_48x48x48x96_pmc.md")
    try:
        txt = open(path).read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", txt)
    return float(m.group(1)) if m else None


def runtime_copy_rate(nbytes=4 << 30, reps=10):
    """The HIP runtime's own device-to-device copy (hipMemcpyAsync of a
    4 GiB buffer), bytes read plus bytes written per second: a streaming
    ceiling measured with code that is not this project's."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)):
        return None
    try:
        hip.hipMemset(src, 1, C.c_size_t(nbytes))
        e0, e1 = C.c_void_p(), C.c_void_p()
        hip.hipEventCreate(C.byref(e0))
        hip.hipEventCreate(C.byref(e1))
        for _ in range(2):
            hip.hipMemcpyAsync(dst, src, C.c_size_t(nbytes), 3, None)
        hip.hipEventRecord(e0, None)
        for _ in range(reps):
            hip.hipMemcpyAsync(dst, src, C.c_size_t(nbytes), 3, None)
        hip.hipEventRecord(e1, None)
        hip.hipEventSynchronize(e1)
        ms = C.c_float()
        hip.hipEventElapsedTime(C.byref(ms), e0, e1)
        per = ms.value / reps
        gbs = 2 * nbytes / (per * 1e-3) / 1e9
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
        return {"bytes": nbytes, "ms": round(per, 4), "gbs_read_plus_write": round(gbs, 1),
                "frac_of_peak": round(gbs / HBM_PEAK_GBS, 4)}
    finally:
        hip.hipFree(src)
        hip.hipFree(dst)


def cpu_baseline(seconds=12.0):
    """The oracle's FP64 clover Dslash (hop + A^-1) timed on the host on a
    16^3x32 sample with its OpenMP site loops on all the threads OpenMP is
    given (OMP_NUM_THREADS), in the same algorithmic GB/s."""
    import numpy as np  # noqa: F401
    import oracle
    L = [16, 16, 16, 32]
    vh = L[0] * L[1] * L[2] * L[3] // 2
    u = oracle.gen_gauge(L, 3, 0.4)
    clov = oracle.clover_build(L, u, CSW, MASS)
    inv = oracle.clover_invert(clov[:vh * 72])
    psi = oracle.gen_spinor(L, 4, 1)
    hp = oracle.hop_params(-1.0, 1.0, 1.0)
    n = 0
    t0 = time.perf_counter()
    while True:
        oracle.clover_apply(inv, oracle.hop(L, u, psi, 1, 0, hp))
        n += 1
        dt = time.perf_counter() - t0
        if dt > seconds:
            break
    per = dt / n
    threads = oracle.lib().orc_num_threads()
    return {"value": round(BYTES_SITE[("f64", False)] * vh / per / 1e9, 3), "unit": "GB/s",
            "cores": threads, "kind": "port", "precision": "FP64, full gauge links",
            "lattice": L, "host": "this run's host cores",
            "sample": f"oracle FP64 clover Dslash (hop + A^-1) on {threads} OpenMP threads, "
                      f"16^3x32, {vh} sites/parity, "
                      f"{n} applications in {dt:.1f} s ({per * 1e3:.1f} ms each, "
                      f"{FLOPS_SITE_CLOVER * vh / per / 1e9:.2f} GFLOP/s)",
            # the reference's own vectorised CPU build cannot run here (its
            # sources stay off this machine); its figure for the headline
            # workload, measured when the reference was surveyed
            "reference_build": {
                "value": round(BYTES_SITE[("f32", True)] * LATT[0] * LATT[1] * LATT[2] * LATT[3] // 2
                               / 91.4e-3 / 1e9, 1),
                "unit": "GB/s", "cores": 8, "kind": "reference",
                "precision": "FP32, two-row gauge links (AVX-512, VECLEN 16, SOALEN 8)",
                "lattice": LATT, "ms_per_dslash": 91.4, "gflops": 106,
                "host": "8-core Xeon @ 2.70 GHz with AVX-512 (the survey container), not this run's host",
                "sample": "clover Dslash A^-1 D psi, one parity, isign +1, median of 3 repeats"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-block", action="store_true")
    ap.add_argument("--no-family", action="store_true",
                    help="skip the Wilson and twisted-mass clover Dslash and the local-volume "
                         "(t-split) measurements")
    ap.add_argument("--no-multi", action="store_true",
                    help="skip the 12-RHS launches (they share the headline kernel's symbol)")
    ap.add_argument("--rhs-tile", default=None,
                    help="reg,rt: fixed multi-RHS tiling (register tile, sources per launch) instead of "
                         "autotuning (profiler runs)")
    ap.add_argument("--rhs-traversal", default=None,
                    help="block,bz,by,lds: fixed traversal of the multi-RHS launches (profiler runs)")
    ap.add_argument("--hop-config", default=None,
                    help="block,bz,slabs: fixed hop-kernel launch configuration instead of "
                         "autotuning (used for profiler runs so every launch is identical)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path is "
              "single-GPU with in-kernel periodic boundaries)", file=sys.stderr)
        sys.exit(2)

    import qphix_amd as q

    if args.hop_config:
        blk, bz, sl = (int(v) for v in args.hop_config.split(","))
        q.call("qphix_tune_config", 2, blk, bz, sl)
    if args.rhs_tile:
        reg, rt = (int(v) for v in args.rhs_tile.split(","))
        q.call("qphix_tune_rhs", reg, rt)
        RHS_TILE[:] = [reg, rt]
    vh = LATT[0] * LATT[1] * LATT[2] * LATT[3] // 2
    t_setup = time.perf_counter()
    g64 = q.Geometry(LATT, q.F64, False)
    g32 = q.Geometry(LATT, q.F32, True)
    g32u = q.Geometry(LATT, q.F32, False)
    u64 = q.gen_gauge(g64, 3, 0.4)
    # clover A on cb 1, A^-1 on cb 0 (what the even-odd operator uses for target cb 1)
    inv64 = q.clover_build(g64, u64, CSW, MASS, 0, q.CLOV_AINV)
    u32 = q.convert_gauge(g32, g64, u64)
    u32u = q.convert_gauge(g32u, g64, u64)
    inv32 = q.convert_clover(g32, g64, inv64)
    inv32u = q.convert_clover(g32u, g64, inv64)
    psi = q.gen_spinor(g32, 5, 1)
    out = g32.spinor()
    psiu = q.gen_spinor(g32u, 5, 1)
    outu = g32u.spinor()
    q.sync()
    log(f"setup {time.perf_counter() - t_setup:.1f} s")

    D = q.Dslash(g32, t_boundary=-1.0)
    Du = q.Dslash(g32u, t_boundary=-1.0)
    timer = q.Timer()

    def run(Dx, o, p, u, inv, k):
        for _ in range(k):
            Dx.clov_dslash(o, p, u[0], inv, 1, 0)

    run(D, out, psi, u32, inv32, args.warmup)
    q.sync()
    t0 = time.perf_counter()
    timer.start()
    run(D, out, psi, u32, inv32, args.steps)
    ms = timer.stop()
    wall = time.perf_counter() - t0
    per_ms = ms / args.steps
    bytes_launch = BYTES_SITE[("f32", True)] * vh
    gbs = bytes_launch / (per_ms * 1e-3) / 1e9
    log(f"clover dslash f32 c12: {per_ms:.4f} ms/launch, {gbs:.1f} GB/s algorithmic")

    # uncompressed gauge, same kernel family (reported alongside)
    run(Du, outu, psiu, u32u, inv32u, args.warmup)
    timer.start()
    run(Du, outu, psiu, u32u, inv32u, args.steps)
    ms18 = timer.stop() / args.steps
    gbs18 = BYTES_SITE[("f32", False)] * vh / (ms18 * 1e-3) / 1e9
    log(f"clover dslash f32 c18: {ms18:.4f} ms/launch, {gbs18:.1f} GB/s algorithmic")

    # the same per-site traffic without the stencil (one work-item per site
    # streaming 48 words in and 6 out, 60 / 6 for c18): the bandwidth this
    # traffic shape reaches on the device, the practical ceiling of the hop
    ceiling = {}
    for name, wi, wo, nb, hop_ms in (("f32_c12", 48, 6, BYTES_SITE[("f32", True)], per_ms),
                                     ("f32_c18", 60, 6, BYTES_SITE[("f32", False)], ms18)):
        v = ctypes.c_double()
        q.call("qphix_stream_calibrate", g32.h, wi, wo, 20, ctypes.byref(v))
        cg = nb * vh / (v.value * 1e-3) / 1e9
        ceiling[name] = {"ms": round(v.value, 5), "gbs": round(cg, 1), "frac_of_peak": round(cg / HBM_PEAK_GBS, 4),
                         "hop_over_ceiling": round(v.value / hop_ms, 4)}
        log(f"streaming ceiling {name}: {v.value:.4f} ms, {cg:.1f} GB/s; hop at {v.value / hop_ms:.1%} of it")

    # multi-right-hand-side clover Dslash: 12 right-hand sides share every link
    # and clover read; algorithmic bytes 8*G + C + n*(S_in + S_out) per site
    nr = 12
    extra = {}
    if not args.no_multi:
        if args.rhs_traversal:
            blk, bz, by, lds = (int(v) for v in args.rhs_traversal.split(","))
            q.call("qphix_tune_traversal", blk, bz, by, 1, lds)
        extra["multi_rhs_12"] = multi_rhs(q, D, g32, u32, inv32, nr, max(1, args.steps // 10),
                                          per_ms, timer)
        if args.rhs_traversal:
            q.call("qphix_tune_config", 1, 256, 0, 1)

    if not args.no_family:
        extra["kernels_48x48x48x96"] = kernel_family(q, g64, u64, g32, u32, psi, out, vh,
                                                     max(10, args.steps // 5))
        extra["local_volumes_t_split"] = local_volumes(q, max(20, args.steps // 2))

    # copy kernel of known size (FP32 spinor, 96 B/site read + 96 B/site written)
    # for the profiler's FETCH_SIZE / WRITE_SIZE calibration; not timed
    for _ in range(5):
        q.copy(g32, out, psi)
    q.sync()

    rc = runtime_copy_rate()
    if rc:
        ceiling["hip_runtime_d2d_copy"] = rc
        log(f"hipMemcpy device to device, 4 GiB: {rc['ms']:.3f} ms, {rc['gbs_read_plus_write']:.1f} GB/s "
            f"({rc['frac_of_peak']:.1%} of peak)")
    extra["streaming_ceiling"] = ceiling
    extra.update({"wall_ms_per_step": round(wall * 1e3 / args.steps, 5),
                  "tuned": q.tune_summary().strip().splitlines(),
                  "ms_per_dslash_c12": round(per_ms, 5), "ms_per_dslash_c18": round(ms18, 5),
                  "gbs_c18": round(gbs18, 1),
                  "gflops_c12": round(FLOPS_SITE_CLOVER * vh / (per_ms * 1e-3) / 1e9, 1)})

    if not args.no_solve:
        extra["solve"] = solve_compare(q, g64, u64)
    if not args.no_block:
        extra["clover_32x32x32x64"] = clover_32(q, max(20, args.steps))
        extra["multi_rhs_32x32x32x64"] = block_compare(q)
    if not args.no_batch:
        extra["batch_tmclover"] = batch_compare(q)
        extra["doublet_tmclover_operator"] = doublet_operator(q)

    result = {
        "metric": "clover Dslash GB/s (% HBM peak); solve s to 1e-10, mixed vs FP64",
        "value": round(gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(per_ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded U = reunit(exp(0.4 i A)) gauge, Gaussian source, "
                "clover c_sw=1 built on the GPU in FP64",
        "config": {"workload": "clover Dslash A^-1 D psi, one parity, 48^3x96, FP32 storage and "
                               "arithmetic, two-row gauge compression, t antiperiodic",
                   "lattice": LATT, "sites_per_parity": vh, "kappa": KAPPA, "c_sw": CSW,
                   "bytes_per_site": BYTES_SITE[("f32", True)]},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profiles(),
                     "traffic_source": "profiles/clover_dslash_f32c12_48x48x48x96_pmc.md (rocprofv3 "
                                       "FETCH_SIZE + WRITE_SIZE, copy-calibrated; recorded with "
                                       "--hop-config 256,6,1, not measured in this run)"},
        "extra": extra,
    }
    if not args.no_cpu:
        log("cpu baseline (oracle, OpenMP threads)...")
        result["cpu_baseline"] = cpu_baseline()
    print(json.dumps(result), flush=True)


def multi_rhs(q, D, g32, u32, inv32, nr, reps, per_ms, timer):
    """Multi-right-hand-side clover Dslash: nr right-hand sides share every link
    and clover read; algorithmic bytes 8*G + C + nr*(S_in + S_out) per site."""
    vh = g32.vh
    psis = [q.gen_spinor(g32, 100 + r, 1) for r in range(nr)]
    outs = [g32.spinor() for _ in range(nr)]
    D.clov_dslash_multi(outs, psis, u32[0], inv32, 1, 0)
    q.sync()
    timer.start()
    for _ in range(reps):
        D.clov_dslash_multi(outs, psis, u32[0], inv32, 1, 0)
    ms_multi = timer.stop() / reps
    bytes_multi = (8 * 12 + 72 + nr * 48) * 4 * vh
    gbs = bytes_multi / (ms_multi * 1e-3) / 1e9
    log(f"clover dslash f32 c12, {nr} rhs: {ms_multi:.3f} ms ({ms_multi / nr:.4f} ms per rhs), "
        f"{gbs:.1f} GB/s algorithmic")
    nb = 2  # the multi-RHS tuning class of nr sources
    while nb * 2 <= nr and nb < 32:
        nb *= 2
    out = {"tuned": [ln for ln in q.tune_summary().splitlines() if f" n{nb}:" in ln],
           "ms": round(ms_multi, 4), "ms_per_rhs": round(ms_multi / nr, 5),
           "gbs_algorithmic": round(gbs, 1),
           "speedup_per_rhs_vs_single": round(per_ms / (ms_multi / nr), 3)}
    # the matrix-core kernel on 16 right-hand sides: each link multiplies the
    # 3 x 32 block of projected half spinors as a real block-diagonal 12 x 12
    # by 12 x 16 product (two sites per v_mfma_f32_16x16x4f32 triple)
    nm = 16
    psis = [q.gen_spinor(g32, 100 + r, 1) for r in range(nm)]
    outs = [g32.spinor() for _ in range(nm)]
    q.call("qphix_tune_rhs", 8, nm)
    D.clov_dslash_multi(outs, psis, u32[0], inv32, 1, 0)
    q.sync()
    timer.start()
    for _ in range(reps):
        D.clov_dslash_multi(outs, psis, u32[0], inv32, 1, 0)
    ms_m = timer.stop() / reps
    q.call("qphix_tune_rhs", *(RHS_TILE or (0, 0)))
    issued = 3072.0 * vh * nm   # MFMA flops issued per launch (incl. the block-diagonal zeros)
    useful = 1152.0 * vh * nm   # 8 hops x 2 spin components x 6 x 6 real multiply-adds
    gbs_m = (8 * 12 + 72 + nm * 48) * 4 * vh / (ms_m * 1e-3) / 1e9
    out["mfma_16rhs"] = {"ms": round(ms_m, 4), "ms_per_rhs": round(ms_m / nm, 5),
                         "gbs_algorithmic": round(gbs_m, 1),
                         "mfma_tflops_issued": round(issued / (ms_m * 1e-3) / 1e12, 2),
                         "mfma_tflops_useful": round(useful / (ms_m * 1e-3) / 1e12, 2),
                         "mfma_utilisation_of_fp32_peak": round(issued / (ms_m * 1e-3) / 157.3e12, 4)}
    log(f"clover dslash f32 c12, {nm} rhs on matrix cores: {ms_m:.3f} ms ({ms_m / nm:.4f} ms per rhs), "
        f"MFMA {issued / (ms_m * 1e-3) / 1e12:.2f} TFLOP/s issued")
    del psis, outs
    # FP16 storage, 16 right-hand sides: the tuned vector tiles against the
    # matrix-core form on v_mfma_f32_16x16x16f16 (one MFMA per site pair and
    # hop: 8192 flops issued for 2 sites x 8 sources, 2304 of them useful)
    g16 = q.Geometry(g32.latt, q.F16, True)
    u16 = q.convert_gauge(g16, g32, u32)
    inv16 = q.convert_clover(g16, g32, inv32)
    D16 = q.Dslash(g16, t_boundary=-1.0)
    psis = [q.gen_spinor(g16, 200 + r, 1) for r in range(nm)]
    outs = [g16.spinor() for _ in range(nm)]

    def time16():
        D16.clov_dslash_multi(outs, psis, u16[0], inv16, 1, 0)
        q.sync()
        timer.start()
        for _ in range(reps):
            D16.clov_dslash_multi(outs, psis, u16[0], inv16, 1, 0)
        return timer.stop() / reps

    ms_v16 = time16()
    q.call("qphix_tune_rhs", 8, nm)
    ms_m16 = time16()
    q.call("qphix_tune_rhs", *(RHS_TILE or (0, 0)))
    issued16 = 4096.0 * vh * nm
    b16 = (196 + 148 + nm * 104) * vh
    out["f16_16rhs"] = {
        "vector_tiles": {"ms": round(ms_v16, 4), "ms_per_rhs": round(ms_v16 / nm, 5),
                         "gbs_algorithmic": round(b16 / (ms_v16 * 1e-3) / 1e9, 1)},
        "mfma": {"ms": round(ms_m16, 4), "ms_per_rhs": round(ms_m16 / nm, 5),
                 "gbs_algorithmic": round(b16 / (ms_m16 * 1e-3) / 1e9, 1),
                 "mfma_tflops_issued": round(issued16 / (ms_m16 * 1e-3) / 1e12, 2),
                 "mfma_tflops_useful": round(useful / (ms_m16 * 1e-3) / 1e12, 2),
                 "mfma_utilisation_of_f16_peak": round(issued16 / (ms_m16 * 1e-3) / 2.5e15, 5)}}
    log(f"clover dslash f16 c12, {nm} rhs: vector tiles {ms_v16:.3f} ms, matrix cores {ms_m16:.3f} ms "
        f"(MFMA {issued16 / (ms_m16 * 1e-3) / 1e12:.2f} TFLOP/s issued)")
    return out


def kernel_family(q, g64, u64, g32, u32, psi, out, vh, reps):
    """The FP32 two-row Wilson Dslash (576 B/site) and twisted-mass clover
    Dslash with full inverse blocks (1152 B/site, mu = 0.01) on the same
    48^3x96 fields, each beside the streaming calibration of its traffic."""
    mu = 0.01
    full64 = q.clover_build(g64, u64, CSW, MASS, 0, q.CLOV_TMINV, mu)
    full32 = q.convert_clover(g32, g64, full64, full=True)
    del full64
    D = q.Dslash(g32, t_boundary=-1.0, prec_mass_rho=mu)
    timer = q.Timer()
    res = {}
    cases = (("wilson_f32c12", 576, 30, lambda: D.dslash(out, psi, u32[0], 1, 0)),
             ("tmclover_f32c12_full_blocks", 1152, 66,
              lambda: D.tmclov_dslash(out, psi, u32[0], [full32, full32], 1, 0, 0)))
    for name, nb, words, launch in cases:
        for _ in range(5):
            launch()
        timer.start()
        for _ in range(reps):
            launch()
        ms = timer.stop() / reps
        c = ctypes.c_double()
        q.call("qphix_stream_calibrate", g32.h, words, 6, 20, ctypes.byref(c))
        gb = nb * vh / (ms * 1e-3) / 1e9
        res[name] = {"ms": round(ms, 5), "bytes_per_site": nb, "gbs_algorithmic": round(gb, 1),
                     "frac_of_peak": round(gb / HBM_PEAK_GBS, 4), "streaming_ceiling_ms": round(c.value, 5),
                     "hop_over_ceiling": round(c.value / ms, 4)}
        log(f"{name}: {ms:.4f} ms, {gb:.1f} GB/s ({gb / HBM_PEAK_GBS:.1%} of peak), "
            f"{c.value / ms:.1%} of its streaming ceiling")
    return res


def local_volumes(q, reps):
    """The FP32 two-row clover Dslash on one GPU's share of 48^3x96 split 2, 4
    and 8 ways in t (48^3x48, 48^3x24, 48^3x12), periodic wrap standing in for
    the halo exchange: the per-GPU kernel of a time-sharded run."""
    res = {}
    timer = q.Timer()
    for nt in (48, 24, 12):
        L = [48, 48, 48, nt]
        g64 = q.Geometry(L, q.F64, False)
        u64 = q.gen_gauge(g64, 3, 0.4)
        inv64 = q.clover_build(g64, u64, CSW, MASS, 0, q.CLOV_AINV)
        g = q.Geometry(L, q.F32, True)
        u = q.convert_gauge(g, g64, u64)
        inv = q.convert_clover(g, g64, inv64)
        del u64, inv64
        psi = q.gen_spinor(g, 5, 1)
        out = g.spinor()
        D = q.Dslash(g, t_boundary=1.0)
        for _ in range(5):
            D.clov_dslash(out, psi, u[0], inv, 1, 0)
        timer.start()
        for _ in range(reps):
            D.clov_dslash(out, psi, u[0], inv, 1, 0)
        ms = timer.stop() / reps
        gb = BYTES_SITE[("f32", True)] * g.vh / (ms * 1e-3) / 1e9
        res[f"48x48x48x{nt}"] = {"split": 96 // nt, "sites_per_parity": g.vh, "ms": round(ms, 5),
                                 "gbs_algorithmic": round(gb, 1), "frac_of_peak": round(gb / HBM_PEAK_GBS, 4)}
        log(f"local volume 48^3x{nt} ({96 // nt}-way t split): {ms:.4f} ms, {gb:.1f} GB/s ({gb / HBM_PEAK_GBS:.1%})")
    return res


def doublet_operator(q, reps=50):
    """Non-degenerate twisted-mass clover doublet Schur operator on 16^3x32
    (mu = 0.01, eps = 0.005, kappa = 0.1240, c_sw = 1): four Wilson hops, the
    two-flavour inverse clover term and the two-flavour A chi - D psi, per
    application, in FP64 and FP32."""
    L = [16, 16, 16, 32]
    mu, eps = 0.01, 0.005
    mass = 1.0 / (2 * 0.1240) - 4.0
    g64 = q.Geometry(L, q.F64, False)
    u64 = q.gen_gauge(g64, 9, 0.4)
    A = q.clover_build(g64, u64, CSW, mass, 1, q.CLOV_A)
    up = q.clover_build(g64, u64, CSW, mass, 0, q.CLOV_ND_UP, mu, eps)
    dn = q.clover_build(g64, u64, CSW, mass, 0, q.CLOV_ND_DN, mu, eps)
    off = q.clover_build(g64, u64, CSW, mass, 0, q.CLOV_ND_OFF, mu, eps)
    timer = q.Timer()
    out = {"lattice": L, "mu": mu, "epsilon": eps}
    for name, prec in (("f64", q.F64), ("f32", q.F32)):
        if prec == q.F64:
            g, u, fA, fup, fdn, foff = g64, u64, A, up, dn, off
        else:
            g = q.Geometry(L, prec, False)
            u = q.convert_gauge(g, g64, u64)
            fA = q.convert_clover(g, g64, A)
            fup = q.convert_clover(g, g64, up, full=True)
            fdn = q.convert_clover(g, g64, dn, full=True)
            foff = q.convert_clover(g, g64, off)
        M = q.Operator(g, q.OP_NDTMCLOVER, u, clov=[None, fA], fullclov=[[fup, fdn], None],
                       clov_odiag=[foff, None], twisted_mass=mu, epsilon=eps, t_boundary=-1.0)
        x = [q.gen_spinor(g, 11, 1), q.gen_spinor(g, 12, 1)]
        y = [g.spinor(), g.spinor()]
        for _ in range(5):
            M(y, x, 1, 1)
        timer.start()
        for _ in range(reps):
            M(y, x, 1, 1)
        ms = timer.stop() / reps
        out[name] = {"ms_per_application": round(ms, 4),
                     "flavour_sites_per_s": round(2 * g.vh / (ms * 1e-3), 1)}
        log(f"doublet tm-clover operator {name} 16^3x32: {ms:.4f} ms per application")
    return out


def block_compare(q):
    """Batched right-hand sides on the 32^3x64 clover operator (kappa = 0.1240,
    c_sw = 1): 12 point sources (spin x colour at one seeded site) and 16
    Gaussian sources solved on M^+M in FP32 with FP64 reductions, by block CG
    and by lockstep CG (one Krylov space per source), to 1e-6; then multishift
    CG with 12 shifts sigma_k = 10^(-4 + 4k/11) in FP64 to 1e-8 per shift.
    True residuals are recomputed with the FP64 operator."""
    import ctypes as C
    import random
    L = [32, 32, 32, 64]
    mass = 1.0 / (2 * 0.1240) - 4.0
    g64 = q.Geometry(L, q.F64, False)
    u64 = q.gen_gauge(g64, 2, 0.4)
    A = q.clover_build(g64, u64, CSW, mass, 1, q.CLOV_A)
    Ai = q.clover_build(g64, u64, CSW, mass, 0, q.CLOV_AINV)
    M64 = q.Operator(g64, q.OP_CLOVER, u64, clov=[None, A], invclov=[Ai, None], t_boundary=-1.0)
    g32 = q.Geometry(L, q.F32, True)
    u32 = q.convert_gauge(g32, g64, u64)
    A32, Ai32 = q.convert_clover(g32, g64, A), q.convert_clover(g32, g64, Ai)
    M32 = q.Operator(g32, q.OP_CLOVER, u32, clov=[None, A32], invclov=[Ai32, None], t_boundary=-1.0)
    vh = g32.vh
    site = random.Random(12).randrange(vh)
    e = np.zeros(vh * 24)
    rhs64 = []
    for k in range(12):
        e[site * 24 + 2 * k] = 1.0
        rhs64.append(g64.spinor().upload(e, 1))
        e[site * 24 + 2 * k] = 0.0
    rhs64 += [q.gen_spinor(g64, 500 + k, 1) for k in range(16)]
    rhs32 = []
    for b in rhs64:
        f = g32.spinor()
        q.call("qphix_convert_spinor", g32.h, f.p, C.c_double(1.0), g64.h, b.p)
        rhs32.append(f)
    n = len(rhs32)
    tol = 1e-6

    def true_resid(xs32):
        worst = 0.0
        x64, t1, t2 = g64.spinor(), g64.spinor(), g64.spinor()
        for x, b in zip(xs32, rhs64):
            q.call("qphix_convert_spinor", g64.h, x64.p, C.c_double(1.0), g32.h, x.p)
            M64(t1, x64, 1, 1)
            M64(t2, t1, -1, 1)
            rn, bn = C.c_double(), C.c_double()
            q.call("qphix_xmy_norm2", g64.h, t2.p, b.p, t2.p, C.byref(rn))
            q.call("qphix_norm2_spinor", g64.h, b.p, C.byref(bn))
            worst = max(worst, (rn.value / bn.value) ** 0.5)
        return worst

    out = {"lattice": L, "n_rhs": n, "tol": tol}
    w = [g32.spinor() for _ in range(n)]
    M32.apply_multi(w, rhs32, 1, 1)  # untimed: tunes the multi-RHS kernels (both signs)
    M32.apply_multi(w, rhs32, -1, 1)
    q.sync()
    xb = [g32.spinor() for _ in range(n)]
    t0 = time.perf_counter()
    stb = q.InvBlockCG(M32, 3000).solve_multi(xb, rhs32, tol)
    q.sync()
    tb = time.perf_counter() - t0
    out["block_cg"] = {"seconds": round(tb, 3), "block_iters": stb[0].n_iters,
                       "status": max(st.status for st in stb), "worst_true_rel_resid": true_resid(xb)}
    xl = [g32.spinor() for _ in range(n)]
    t0 = time.perf_counter()
    stl = q.InvCG(M32, 5000).solve_multi(xl, rhs32, tol)
    q.sync()
    tl = time.perf_counter() - t0
    out["lockstep_cg"] = {"seconds": round(tl, 3), "max_iters": max(st.n_iters for st in stl),
                          "status": max(st.status for st in stl), "worst_true_rel_resid": true_resid(xl)}
    log(f"block cg 32^3x64 x{n}: {tb:.2f} s ({stb[0].n_iters} block iterations), lockstep cg "
        f"{tl:.2f} s (max {out['lockstep_cg']['max_iters']} iterations)")
    del xb, xl, w, rhs32
    # multishift CG, one Gaussian source, FP64: scalars and recurrences on
    # the host (one round trip per iteration), then in device memory with the
    # iteration replayed from a HIP graph (the library's default)
    shifts = [10 ** (-4 + 4 * k / 11) for k in range(12)]
    b = rhs64[12]
    per_mode = {}
    for mode in (0, 1, 2):
        xs = [g64.spinor() for _ in shifts]
        ms = q.MInvCG(M64, 5000, len(shifts), mode=mode)
        q.sync()
        t0 = time.perf_counter()
        nit, fin, mv = ms(xs, b, shifts, [1e-8] * len(shifts))
        q.sync()
        per_mode[mode] = (time.perf_counter() - t0, nit)
    tm = per_mode[2][0]
    t1, t2 = g64.spinor(), g64.spinor()
    worst = 0.0
    for sg, x in zip(shifts, xs):
        M64(t1, x, 1, 1)
        M64(t2, t1, -1, 1)
        q.call("qphix_axpy", g64.h, C.c_double(sg), x.p, t2.p)
        rn, bn = C.c_double(), C.c_double()
        q.call("qphix_xmy_norm2", g64.h, t2.p, b.p, t2.p, C.byref(rn))
        q.call("qphix_norm2_spinor", g64.h, b.p, C.byref(bn))
        worst = max(worst, (rn.value / bn.value) ** 0.5)
    out["multishift_cg"] = {"shifts": len(shifts), "seconds": round(tm, 4), "iters": nit,
                            "mv_apps": mv, "worst_true_rel_resid_over_shifts": worst,
                            "seconds_host_scalars": round(per_mode[0][0], 4),
                            "seconds_device_scalars_eager": round(per_mode[1][0], 4),
                            "seconds_device_scalars_graph": round(per_mode[2][0], 4)}
    log(f"multishift cg 32^3x64, 12 shifts: {tm:.3f} s ({nit} iterations; host scalars "
        f"{per_mode[0][0]:.3f} s, device scalars eager {per_mode[1][0]:.3f} s), worst true resid {worst:.2e}")
    return out


def clover_32(q, reps):
    """The 32^3x64 clover configuration (kappa = 0.1240, c_sw = 1, FP32 fields,
    FP64 reductions, t antiperiodic): the single-parity clover Dslash
    A^-1 D psi in FP32 with two-row and with full gauge links against the HBM
    roofline, and BiCGStab on the even-odd clover operator (odd target
    checkerboard, A on odd sites, A^-1 on even sites) to 1e-7, its true
    residual recomputed with the FP64 operator of the same seeded fields."""
    import ctypes as C
    L = [32, 32, 32, 64]
    mass = 1.0 / (2 * 0.1240) - 4.0
    g64 = q.Geometry(L, q.F64, False)
    u64 = q.gen_gauge(g64, 2, 0.4)
    A = q.clover_build(g64, u64, CSW, mass, 1, q.CLOV_A)
    Ai = q.clover_build(g64, u64, CSW, mass, 0, q.CLOV_AINV)
    M64 = q.Operator(g64, q.OP_CLOVER, u64, clov=[None, A], invclov=[Ai, None], t_boundary=-1.0)
    out = {"lattice": L, "kappa": 0.1240, "c_sw": CSW, "sites_per_parity": g64.vh}
    timer = q.Timer()
    for c12 in (True, False):
        g = q.Geometry(L, q.F32, c12)
        u = q.convert_gauge(g, g64, u64)
        inv = q.convert_clover(g, g64, Ai)
        psi, o = q.gen_spinor(g, 9, 1), g.spinor()
        D = q.Dslash(g, t_boundary=-1.0)
        for _ in range(5):
            D.clov_dslash(o, psi, u[0], inv, 1, 0)
        timer.start()
        for _ in range(reps):
            D.clov_dslash(o, psi, u[0], inv, 1, 0)
        ms = timer.stop() / reps
        nb = BYTES_SITE[("f32", c12)]
        gbs = nb * g.vh / (ms * 1e-3) / 1e9
        out["dslash_f32_c12" if c12 else "dslash_f32_c18"] = {
            "ms": round(ms, 5), "bytes_per_site": nb, "gbs_algorithmic": round(gbs, 1),
            "frac_of_peak": round(gbs / HBM_PEAK_GBS, 4),
            "gflops": round(FLOPS_SITE_CLOVER * g.vh / (ms * 1e-3) / 1e9, 1)}
        log(f"32^3x64 clover dslash f32 {'c12' if c12 else 'c18'}: {ms:.4f} ms, {gbs:.1f} GB/s "
            f"({gbs / HBM_PEAK_GBS:.1%} of peak)")
        if not c12:
            continue
        A32 = q.convert_clover(g, g64, A)
        M32 = q.Operator(g, q.OP_CLOVER, u, clov=[None, A32], invclov=[inv, None], t_boundary=-1.0)
        b64 = q.gen_spinor(g64, 8, 1)
        b32 = g.spinor()
        q.call("qphix_convert_spinor", g.h, b32.p, C.c_double(1.0), g64.h, b64.p)
        w = g.spinor()
        M32(w, b32, 1, 1)  # untimed: tunes the hop kernels of the operator
        x = g.spinor()
        s = q.InvBiCGStab(M32, 5000)
        q.sync()
        t0 = time.perf_counter()
        st = s(x, b32, 1e-7)
        q.sync()
        ts = time.perf_counter() - t0
        x64, r = g64.spinor(), g64.spinor()
        q.call("qphix_convert_spinor", g64.h, x64.p, C.c_double(1.0), g.h, x.p)
        M64(r, x64, 1, 1)
        rn, bn = C.c_double(), C.c_double()
        q.call("qphix_xmy_norm2", g64.h, r.p, b64.p, r.p, C.byref(rn))
        q.call("qphix_norm2_spinor", g64.h, b64.p, C.byref(bn))
        tr = float((rn.value / bn.value) ** 0.5)
        out["bicgstab_f32"] = {"seconds": round(ts, 4), "iters": st.n_iters, "mv_apps": st.mv_apps,
                               "site_flops": st.site_flops, "status": st.status, "rsd_target": 1e-7,
                               "true_rel_resid_fp64": tr}
        log(f"32^3x64 bicgstab f32: {ts:.3f} s, {st.n_iters} iterations, true residual (FP64) {tr:.2e}")
    return out


def solve_compare(q, g64, u64):
    """Clover solve on the target checkerboard to true residual 1e-10:
    pure-FP64 BiCGStab vs FP64 defect correction around an FP16-storage
    BiCGStab (inner relaxation 1e-2)."""
    import ctypes as C
    g16 = q.Geometry(LATT, q.F16, True)
    clov64 = [None, q.clover_build(g64, u64, CSW, MASS, 1, q.CLOV_A)]
    inv64 = [q.clover_build(g64, u64, CSW, MASS, 0, q.CLOV_AINV), None]
    u16 = q.convert_gauge(g16, g64, u64)
    clov16 = [None, q.convert_clover(g16, g64, clov64[1])]
    inv16 = [q.convert_clover(g16, g64, inv64[0]), None]
    M64 = q.Operator(g64, q.OP_CLOVER, u64, clov=clov64, invclov=inv64, t_boundary=-1.0)
    M16 = q.Operator(g16, q.OP_CLOVER, u16, clov=clov16, invclov=inv16, t_boundary=-1.0)
    b = q.gen_spinor(g64, 7, 1)
    out = {}
    # untimed first applications: the hop kernels of both precisions are
    # autotuned on first use
    w64, w16 = g64.spinor(), g16.spinor()
    M64(w64, b, 1, 1)
    M16(w16, g16.spinor(), 1, 1)
    q.sync()
    # pure FP64 BiCGStab
    x64 = g64.spinor()
    s64 = q.InvBiCGStab(M64, 5000)
    q.sync()
    t0 = time.perf_counter()
    st = s64(x64, b, 1e-10)
    q.sync()
    t64 = time.perf_counter() - t0
    out["fp64_bicgstab"] = {"seconds": round(t64, 3), "iters": st.n_iters, "status": st.status}
    # mixed: FP64 outer Richardson, FP16-storage BiCGStab inner
    xm = g64.spinor()
    inner = q.InvBiCGStab(M16, 5000)
    sm = q.InvRichardsonMultiPrec(M64, inner, 1e-2, 200)
    q.sync()
    t0 = time.perf_counter()
    stm = sm(xm, b, 1e-10)
    q.sync()
    tm = time.perf_counter() - t0
    # true residuals and the difference of the solutions, in FP64
    r = g64.spinor()
    M64(r, xm, 1, 1)
    rn, bn = C.c_double(), C.c_double()
    q.call("qphix_xmy_norm2", g64.h, r.p, b.p, r.p, C.byref(rn))
    q.call("qphix_norm2_spinor", g64.h, b.p, C.byref(bn))
    d = g64.spinor()
    dn, xn = C.c_double(), C.c_double()
    q.call("qphix_xmy_norm2", g64.h, d.p, xm.p, x64.p, C.byref(dn))
    q.call("qphix_norm2_spinor", g64.h, x64.p, C.byref(xn))
    out["mixed_fp16_richardson"] = {"seconds": round(tm, 3), "outer_iters": stm.n_iters,
                                    "mv_apps": stm.mv_apps, "status": stm.status,
                                    "true_rel_resid": float((rn.value / bn.value) ** 0.5),
                                    "rel_err_vs_fp64": float((dn.value / xn.value) ** 0.5)}
    out["speedup"] = round(t64 / tm, 3)
    log(f"solve: fp64 {t64:.2f} s ({st.n_iters} it), mixed {tm:.2f} s, speedup {t64 / tm:.2f}")
    # the same defect correction around a BF16-storage inner BiCGStab
    gb = q.Geometry(LATT, q.BF16, True)
    ub = q.convert_gauge(gb, g64, u64)
    clovb = [None, q.convert_clover(gb, g64, clov64[1])]
    invb = [q.convert_clover(gb, g64, inv64[0]), None]
    Mb = q.Operator(gb, q.OP_CLOVER, ub, clov=clovb, invclov=invb, t_boundary=-1.0)
    wb = gb.spinor()
    Mb(wb, gb.spinor(), 1, 1)
    xb = g64.spinor()
    sb = q.InvRichardsonMultiPrec(M64, q.InvBiCGStab(Mb, 5000), 1e-2, 200)
    q.sync()
    t0 = time.perf_counter()
    stb = sb(xb, b, 1e-10)
    q.sync()
    tb = time.perf_counter() - t0
    M64(r, xb, 1, 1)
    q.call("qphix_xmy_norm2", g64.h, r.p, b.p, r.p, C.byref(rn))
    q.call("qphix_xmy_norm2", g64.h, d.p, xb.p, x64.p, C.byref(dn))
    out["mixed_bf16_richardson"] = {"seconds": round(tb, 3), "outer_iters": stb.n_iters,
                                    "mv_apps": stb.mv_apps, "status": stb.status,
                                    "true_rel_resid": float((rn.value / bn.value) ** 0.5),
                                    "rel_err_vs_fp64": float((dn.value / xn.value) ** 0.5),
                                    "speedup": round(t64 / tb, 3)}
    log(f"solve: mixed with bf16 inner {tb:.3f} s ({stb.n_iters} outer), speedup {t64 / tb:.2f}")
    # the clover Dslash A^-1 D psi of each precision the solve uses
    timer = q.Timer()
    per = {}
    for name, g, u, inv, nb in (("f64_c18", g64, u64, inv64[0], BYTES_SITE[("f64", False)]),
                                ("f16_c12", g16, u16, inv16[0], BYTES_SITE[("f16", True)]),
                                ("bf16_c12", gb, ub, invb[0], BYTES_SITE[("bf16", True)])):
        Dp = q.Dslash(g, t_boundary=-1.0)
        ps, o = q.gen_spinor(g, 5, 1), g.spinor()
        for _ in range(5):
            Dp.clov_dslash(o, ps, u[0], inv, 1, 0)
        timer.start()
        for _ in range(20):
            Dp.clov_dslash(o, ps, u[0], inv, 1, 0)
        ms = timer.stop() / 20
        gbs = nb * g.vh / (ms * 1e-3) / 1e9
        per[name] = {"ms": round(ms, 4), "bytes_per_site": nb, "gbs_algorithmic": round(gbs, 1),
                     "frac_of_peak": round(gbs / HBM_PEAK_GBS, 4)}
        log(f"clover dslash {name}: {ms:.4f} ms, {gbs:.1f} GB/s algorithmic")
    out["dslash_by_precision"] = per
    return out


def batch_compare(q):
    """Uneven batch of twisted-mass clover solves: lattices 8^4, 12^3x24,
    16^3x32, 24^3x48 (own gauge seed, mu ~ U[0.001, 0.01]), 4 Gaussian sources
    each; FP16-storage inner CG on M^+M, FP64 outer refinement to 1e-10.
    Batched (one launch per step for all 16 systems, graph-replayed CG
    iterations) against solving the systems one by one."""
    import random
    lats = [[8, 8, 8, 8], [12, 12, 12, 24], [16, 16, 16, 32], [24, 24, 24, 48]]
    mass = 1.0 / (2 * 0.1240) - 4.0
    rng = random.Random(5)
    keep, ops64, ops16, rhs, gs = [], [], [], [], []
    for i, L in enumerate(lats):
        mu = rng.uniform(0.001, 0.01)
        g64 = q.Geometry(L, q.F64, False)
        g16 = q.Geometry(L, q.F16, True)
        u = q.gen_gauge(g64, 5 + i, 0.4)
        A = q.clover_build(g64, u, CSW, mass, 1, q.CLOV_A)
        fp = q.clover_build(g64, u, CSW, mass, 0, q.CLOV_TMINV, mu)
        fm = q.clover_build(g64, u, CSW, mass, 0, q.CLOV_TMINV, -mu)
        u16 = q.convert_gauge(g16, g64, u)
        A16 = q.convert_clover(g16, g64, A)
        fp16 = q.convert_clover(g16, g64, fp, full=True)
        fm16 = q.convert_clover(g16, g64, fm, full=True)
        kw = dict(twisted_mass=mu, prec_mass_rho=mu, t_boundary=-1.0)
        M64 = q.Operator(g64, q.OP_TMCLOVER, u, clov=[None, A], fullclov=[[fp, fm], None], **kw)
        M16 = q.Operator(g16, q.OP_TMCLOVER, u16, clov=[None, A16], fullclov=[[fp16, fm16], None], **kw)
        keep += [g64, g16, u, A, fp, fm, u16, A16, fp16, fm16, M64, M16]
        for r in range(4):
            ops64.append(M64)
            ops16.append(M16)
            gs.append(g64)
            rhs.append(q.gen_spinor(g64, 100 * i + r, 1))
    # warm-up: first applications tune the hop kernels of each lattice
    for M, g in zip(ops64[::4] + ops16[::4], gs[::4] + [o.geom for o in ops16[::4]]):
        a, b = g.spinor(), g.spinor()
        M(a, b, 1, 1)
        M(a, b, -1, 1)
    q.sync()
    B = q.Batch(ops16)
    xb = [g.spinor() for g in gs]
    B.richardson(ops64, xb, rhs, 1e-10, delta=1e-2, graph_steps=8)  # untimed warm-up
    xb = [g.spinor() for g in gs]
    q.sync()
    t0 = time.perf_counter()
    stb = B.richardson(ops64, xb, rhs, 1e-10, delta=1e-2, graph_steps=8)
    q.sync()
    tb = time.perf_counter() - t0
    xs = [g.spinor() for g in gs]
    solvers = []
    for M64, M16 in zip(ops64, ops16):
        inner = q.InvCG(M16, 2000)
        solvers.append(q.InvRichardsonMultiPrec(M64, inner, 1e-2, 50, mdagm=True))
    q.sync()
    t0 = time.perf_counter()
    sts = [sv(x, b, 1e-10) for sv, x, b in zip(solvers, xs, rhs)]
    q.sync()
    ts = time.perf_counter() - t0
    ok_b = all(st.status == q.OK for st in stb)
    ok_s = all(st.status == q.OK for st in sts)
    worst = max(st.rsd_sq_final for st in stb) ** 0.5
    apps_b = sum(int(st.mv_apps) for st in stb)
    apps_s = sum(int(st.mv_apps) for st in sts)
    log(f"batch tm-clover: batched {tb:.3f} s, one-by-one {ts:.3f} s, speedup {ts / tb:.2f}; "
        f"operator applications {apps_b} batched, {apps_s} one by one")
    return {"lattices": lats, "rhs_per_lattice": 4, "batched_seconds": round(tb, 4),
            "one_by_one_seconds": round(ts, 4), "speedup": round(ts / tb, 3),
            "all_converged": ok_b and ok_s, "worst_true_rel_resid_batched": worst,
            "operator_applications_batched": apps_b, "operator_applications_one_by_one": apps_s}


if __name__ == "__main__":
    main()

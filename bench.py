"""Benchmark: fused block-sparse DMRG matvec on the L=100, chi=2048 U(1) Heisenberg chain.

One step = one application of the two-site effective Hamiltonian
L.Wl.Wr.R . theta at the central bond (sites 49, 50) of a U(1) MPS with the
charge profile of a converged chi=2048 state (tenax_amd.workloads), FP64 MFMA.
Reported alongside: the plain per-block GPU path on the same problem, the FP32
and split-FP32 variants with their error, one full 50-step device Lanczos
solve, the dominant kernel's roofline, the oracle's per-combination CPU
matvec (the reference's loop structure) on the same problem, and one full
left-to-right DMRG sweep over the chi = 2048 chain (phases timed, two centre
updates also on the per-block path).

The same line carries two summary keys: "c5", the batched Z2 HOTRG workload
(64 2D-Ising temperatures T = 2.0 + 0.6 k/63, chi = 48, 30 coarse-graining
steps, all instances in lockstep, one batched factorisation per step), and
"c4", the iPEPS CTMRG workload (Heisenberg gate, D = 4, d = 2, chi = 64, QR
projectors, site tensor N(0, 1) from seed 3: CTM sweeps on the dense and the
block-sparse path, the 100-sweep AD forward and one GMRES cycle of its
implicit backward).  ``--workload hotrg|ctm`` runs either alone as the line's
headline (with its CPU baseline).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-sweep] [--no-c4c5]
                       [--workload dmrg|hotrg|ctm]
Prints ONE JSON line on stdout; progress goes to stderr.
"""

import argparse
import importlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DMRG matvec GFLOP/s & sweep s, fused vs per-block GPU at |ΔE|<1e-10"
PEAK_FP64_MFMA_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
GEMM_ONLY = 16  # TNX_STAGE_GEMM_ONLY (include/tnx_dmrg.h)
REDUCE_ONLY = 32  # TNX_STAGE_REDUCE_ONLY


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def build_problem(L, chi, seed):
    import importlib

    import torch

    from tenax_amd.workloads import centre_bond_problem

    pd = importlib.import_module("tenax_amd.algorithms.dmrg")
    t0 = time.time()
    prob, ops, mps, mpo = centre_bond_problem(L, chi, seed)
    torch.cuda.synchronize()
    log(f"setup {time.time() - t0:.1f}s: theta blocks={len(prob.keys)} size={prob.total}")
    return pd, prob, ops, mps, mpo


def committed_traffic(kernel, wgs):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3
    --pmc summary (FETCH_SIZE / WRITE_SIZE, copy-calibrated), or None when the
    summary holds no line for this kernel and launch size."""
    path = os.path.join(ROOT, "profiles", "dmrg_matvec_chi2048_traffic.md")
    try:
        for line in open(path):
            f = line.split()
            if len(f) == 4 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel and int(f[2]) == wgs:
                return float(f[3])
    except OSError:
        pass
    return None


TILE_SHAPES = {0: (64, 64), 1: (128, 64), 2: (64, 128), 3: (128, 128)}


def stage_kernel(prob, stage):
    """Kernel name (as rocprofv3 prints it) and workgroup count of a matvec GEMM stage."""
    t = prob.plan.tiling(stage)
    bm, bn = TILE_SHAPES[t["shape"]]
    lay = 1 if stage == 1 else 0  # stage 1 reads L as K x M, stage 3 reads Y as M x K
    # last two template arguments: fermionic signs (the Heisenberg chain is
    # bosonic) and permuted output stores (the matvec writes plain blocks)
    return f"grouped_gemm_kernel<{lay},0,{bm},{bn},false,false>", t["n_tasks"], t


def time_events(fn, reps, warm=0):
    """Average duration (ms) of fn over reps calls, HIP events on the current
    stream, after warm untimed calls launched back to back with them."""
    import torch

    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    for _ in range(warm):
        fn()
    s.record()
    for _ in range(reps):
        fn()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / reps  # ms


def capture_matvecs(prob, n):
    """One torch.cuda graph holding n consecutive fused matvecs of prob.x
    (recorded on a side stream); returns (graph, output of the last)."""
    import torch

    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        prob.matvec(prob.x, 0)
        torch.cuda.synchronize()
        with torch.cuda.graph(g, stream=s):
            for _ in range(n):
                y = prob.matvec(prob.x, 0)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    return g, y


class FenceFreeEvents:
    """HIP timing events created with hipEventDisableSystemFence: recording one
    between two kernels does not write back and invalidate the caches (a
    default event's system-scope release does, which slows the next kernel by
    10-30 us), so events between the kernels of a matvec time each kernel in
    place."""

    FLAG = 0x20000000  # hipEventDisableSystemFence (hip/hip_runtime_api.h)

    def __init__(self, n):
        import ctypes as C

        self.C = C
        self.hip = C.CDLL("libamdhip64.so.7")  # the runtime torch and libtnx already share
        self.ev = []
        for _ in range(n):
            e = C.c_void_p()
            if self.hip.hipEventCreateWithFlags(C.byref(e), C.c_uint(self.FLAG)) != 0:
                raise RuntimeError("hipEventCreateWithFlags failed")
            self.ev.append(e)

    def record(self, i, stream):
        if self.hip.hipEventRecord(self.ev[i], stream) != 0:
            raise RuntimeError("hipEventRecord failed")

    def ms(self, i, j):
        t = self.C.c_float()
        if self.hip.hipEventElapsedTime(self.C.byref(t), self.ev[i], self.ev[j]) != 0:
            raise RuntimeError("hipEventElapsedTime failed")
        return t.value

    def close(self):
        for e in self.ev:
            self.hip.hipEventDestroy(e)
        self.ev = []


def stage_times(prob, reps, matvec_ms, compute=0):
    """Average duration (ms) of each kernel of the fused matvec inside the
    matvec: coefficients ("0"), stage-1 GEMM and its split-K reduction
    ("1_gemm", "1_reduce"), mix ("2"), stage-3 GEMM and reduction ("3_gemm",
    "3_reduce").

    reps matvecs are launched kernel by kernel on the current stream (after 3
    untimed ones) with a fence-free HIP event before and after every kernel
    ("events": per-kernel averages, "matvec" the first to the last event).
    The events themselves lengthen the matvec by -2% to +12% from machine to
    machine, so each kernel is given its event share of the timed matvec:
    kernel = events[kernel] * matvec_ms / events["matvec"], where matvec_ms
    is the matvec time of the timed steps (replayed from a captured graph,
    no events; "sum" adds the kernels and equals matvec_ms).  "isolated": each kernel launched reps
    times back to back on its own (default events around the reps), which has
    run the GEMMs from 10% faster to 20% slower than inside the matvec."""
    import ctypes as C

    import torch

    from tenax_amd import _lib

    lib = _lib.load()
    L, Wl, Wr, R = prob.envs
    y = torch.empty_like(prob.x)
    stream = _lib.stream_handle()
    args = lambda st: (prob.plan.ptr, st, _lib.dptr(L.data), _lib.dptr(Wl.data), _lib.dptr(Wr.data),
                       _lib.dptr(R.data), _lib.dptr(prob.x), _lib.dptr(y), compute, stream)
    seq = [("0", 0), ("1_gemm", 1 | GEMM_ONLY), ("1_reduce", 1 | REDUCE_ONLY), ("2", 2), ("3_gemm", 3 | GEMM_ONLY),
           ("3_reduce", 3 | REDUCE_ONLY)]
    split = {1: prob.plan.tiling(1)["kchunk"] != 0, 3: prob.plan.tiling(3)["kchunk"] != 0}
    timed = [(name, st) for name, st in seq if not (name.endswith("_reduce") and not split[int(name[0])])]
    run = lambda st: _lib.check(lib.tnx_matvec_execute_stage(*args(st)))
    for _ in range(3):  # populate the intermediates, warm up
        for _, st in seq:
            run(st)
    torch.cuda.synchronize()
    n = len(timed) + 1
    ev = FenceFreeEvents(reps * n)
    try:
        for r in range(reps):
            ev.record(r * n, stream)
            for i, (_, st) in enumerate(timed):
                run(st)
                ev.record(r * n + i + 1, stream)
        torch.cuda.synchronize()
        events = {name: sum(ev.ms(r * n + i, r * n + i + 1) for r in range(reps)) / reps
                  for i, (name, _) in enumerate(timed)}
        events["matvec"] = sum(ev.ms(r * n, r * n + n - 1) for r in range(reps)) / reps
    finally:
        ev.close()
    scale = matvec_ms / events["matvec"]
    out = {name: events.get(name, 0.0) * scale for name, _ in seq}  # no reduction launch for a whole-K tiling
    out["events"] = events
    out["matvec"] = matvec_ms
    out["1"] = out["1_gemm"] + out["1_reduce"]
    out["3"] = out["3_gemm"] + out["3_reduce"]
    out["sum"] = sum(out[name] for name, _ in seq)
    out["isolated"] = {name: time_events(lambda: run(st), reps, warm=3) for name, st in timed}
    for _, st in seq:  # leave complete intermediates behind
        run(st)
    torch.cuda.synchronize()
    f4 = (C.c_double * 4)()
    _lib.check(lib.tnx_fused_stage_flops(prob.plan.ptr, f4))
    return out, list(f4)


def host_info():
    """CPU model, usable cores and the BLAS thread settings of this host."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    env = {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")}
    return {"cpu_model": model, "cpu_count": os.cpu_count(), "usable_cpus": usable, "thread_env": env}


def _trials(fn, n):
    """1 untimed warm-up run, then n timed runs (benchmarks/runner.py:64-117)."""
    fn()
    ts = []
    for _ in range(n):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return ts


def cpu_baseline(prob, trials=3):
    """The oracle's per-block-combination matvec (the reference's loop structure,
    dmrg.py:1048-1151, NumPy FP64 with OpenBLAS) on the same centre-bond problem:
    1 warm-up + `trials` timed runs with all usable cores; the single-thread rate
    from the same runs on a bounded sample (every 8th theta block by size) at 1
    and at all threads, scaled by their ratio."""
    import numpy as np
    from threadpoolctl import threadpool_info, threadpool_limits

    from oracle import core as oc
    from oracle import dmrg as od
    from tests.helpers import to_oracle

    L, Wl, Wr, R = prob.envs
    oL, oWl, oWr, oR = (to_oracle(t) for t in (L, Wl, Wr, R))
    ox = to_oracle(prob.extract(prob.present_after(1)))
    od.MODE["eval"] = "blocks"
    info = host_info()
    n_threads = max((p.get("num_threads") or 0) for p in threadpool_info()) or info["usable_cpus"]
    out = {}
    full = _trials(lambda: out.__setitem__("y", od.matvec(oL, ox, oWl, oWr, oR)), trials)
    y = out.pop("y")
    by_size = sorted(ox.blocks, key=lambda k: -ox.blocks[k].size)
    ox_s = oc.SymT({k: ox.blocks[k] for k in by_size[::8]}, ox.legs, validate=False)
    samp_all = _trials(lambda: od.matvec(oL, ox_s, oWl, oWr, oR), trials)
    with threadpool_limits(limits=1):
        samp_one = _trials(lambda: od.matvec(oL, ox_s, oWl, oWr, oR), trials)
    # agreement with the device result on this problem
    yd = prob.matvec(prob.x, 0)
    from tenax_amd import SymmetricTensor

    yt = to_oracle(SymmetricTensor._wrap(prob.indices, prob.keys, prob.shapes, prob.offsets, yd))
    num = max(float(np.max(np.abs(yt.blocks[k] - b))) for k, b in y.blocks.items())
    den = max(float(np.max(np.abs(b))) for b in y.blocks.values())
    t_full = float(np.median(full))
    ratio = float(np.median(samp_one)) / float(np.median(samp_all))
    return {"seconds_trials": full, "seconds": t_full, "threads": n_threads,
            "sample_seconds_all_threads": samp_all, "sample_seconds_1_thread": samp_one,
            "single_thread_seconds_est": t_full * ratio, "max_rel_diff_vs_gpu": num / den, "host": info}


def centre_envs(pd, mps, mpo):
    """Left and right environments of the centre bond (L/2-1, L/2) of an MPS."""
    L = len(mpo)
    c = L // 2 - 1
    ops = pd._symmetric_ops()
    left = ops.build_trivial_left_env()
    for i in range(c):
        left = ops.update_left_env(left, mps[i], mpo[i])
    right = ops.build_trivial_right_env()
    for i in range(L - 1, c + 1, -1):
        right = ops.update_right_env(right, mps[i], mpo[i])
    return c, left, right


def compare_updates(pd, mps, mpo, chi, reorth, n_bonds=2):
    """Fused vs plain per-block GPU path over the same two-site updates at the
    centre of an MPS (same inputs, no SVD): seconds and Lanczos steps."""
    import torch

    from tenax_amd import DMRGConfig

    L = len(mpo)
    cfg = DMRGConfig(max_bond_dim=chi, num_sweeps=1, lanczos_max_iter=50, lanczos_reorth=reorth)
    ops = pd._symmetric_ops()
    left = pd._build_left_environments_list(mps, mpo, L, ops)
    right = pd._build_right_environments_list(mps, mpo, L, ops)
    c = L // 2 - 1
    cmp = {}
    for name, fn in (("fused", ops.two_site_update), ("per_block", pd._per_block_two_site)):
        fn(mps[c], mps[c + 1], left[c], mpo[c], mpo[c + 1], right[c + 2], cfg)  # warm-up (plans, libraries)
        steps = []
        torch.cuda.synchronize()
        t0 = time.time()
        for i in range(c, c + n_bonds):
            fn(mps[i], mps[i + 1], left[i], mpo[i], mpo[i + 1], right[i + 2], cfg)
            steps.append(pd.LAST_STEPS["n"])
        torch.cuda.synchronize()
        cmp[name] = {"seconds": time.time() - t0, "lanczos_steps": steps}
    cmp["bonds"] = n_bonds
    cmp["speedup_fused"] = cmp["per_block"]["seconds"] / cmp["fused"]["seconds"]
    return cmp


def sweep_comparison(pd, mps_tn, mpo_tn, chi, reorth, fused_only=False):
    """One full sweep (L->R, right environments rebuilt, R->L; dmrg.py:179-265)
    from the same MPS on the fused/grouped path and on the plain per-block GPU
    path (one launch per block product, Lanczos scalars through the host, one
    SVD call per sector): seconds and phases of each."""
    import torch

    from tenax_amd import DMRGConfig, dmrg

    out = {}
    pd.PROGRESS["every"] = 10
    for name, per_block in (("fused", False), ("per_block", True))[: 1 if fused_only else 2]:
        cfg = DMRGConfig(max_bond_dim=chi, num_sweeps=1, lanczos_max_iter=50, lanczos_reorth=reorth,
                         per_block=per_block)
        pd.TIMERS.clear()
        pd.TIMERS["enabled"] = True
        torch.cuda.synchronize()
        t0 = time.time()
        res = dmrg(mpo_tn, mps_tn, cfg)
        torch.cuda.synchronize()
        out[name] = {"seconds": time.time() - t0, "energy": res.energy,
                     "phases_s": {k: v for k, v in pd.TIMERS.items() if k != "enabled"}}
        pd.TIMERS["enabled"] = False
        log(f"sweep {name}: {out[name]['seconds']:.2f}s E={res.energy:.12f} phases {out[name]['phases_s']}")
    pd.PROGRESS["every"] = 0
    if not fused_only:
        out["speedup_fused"] = out["per_block"]["seconds"] / out["fused"]["seconds"]
        out["abs_dE_fused_vs_per_block"] = abs(out["fused"]["energy"] - out["per_block"]["energy"])
    return out


def segment_comparison(pd, mps, mpo, chi, reorth, n_bonds=20):
    """Full left-to-right bond updates (two-site Lanczos, truncated SVD, left
    environment update; dmrg.py:184-214) over the n_bonds bonds around the
    centre, from the same MPS and environments, on the fused/grouped path and
    on the same path with every Lanczos solve replayed from a captured graph
    re-recorded per bond (lanczos_graph), and on the plain per-block GPU path
    (one launch per block product, Lanczos scalars through the host, one SVD
    call per sector)."""
    import torch

    from tenax_amd import DMRGConfig

    L = len(mpo)
    ops = pd._symmetric_ops()
    left0 = pd._build_left_environments_list(mps, mpo, L, ops)
    right = pd._build_right_environments_list(mps, mpo, L, ops)
    c0 = L // 2 - n_bonds // 2
    out = {"bonds": [c0, c0 + n_bonds - 1]}
    for name, per_block, graph in (("fused_eager", False, False), ("fused", False, True), ("per_block", True, False)):
        cfg = DMRGConfig(max_bond_dim=chi, num_sweeps=1, lanczos_max_iter=50, lanczos_reorth=reorth,
                         per_block=per_block, lanczos_graph=graph)
        update = pd._per_block_two_site if per_block else ops.two_site_update
        sites, left = list(mps), list(left0)
        pd._ACTIVE_CONFIG["cfg"] = cfg
        try:
            torch.cuda.synchronize()
            t0 = time.time()
            for i in range(c0, c0 + n_bonds):
                th, e = update(sites[i], sites[i + 1], left[i], mpo[i], mpo[i + 1], right[i + 2], cfg)
                A, s_, B, err = pd._svd_and_truncate_site(th, i, cfg)
                sites[i], sites[i + 1] = A, B
                left[i + 1] = ops.update_left_env(left[i], A, mpo[i])
            torch.cuda.synchronize()
            out[name] = {"seconds": time.time() - t0, "last_energy": float(e)}
        finally:
            pd._ACTIVE_CONFIG.pop("cfg", None)
            pd.release_graphs()
        log(f"segment {name}: {out[name]}")
    out["speedup_fused"] = out["per_block"]["seconds"] / out["fused"]["seconds"]
    return out


def run_c3(pd, L=100, chi=2048, sweeps=6, seed=2, reorth=True, compute="f64"):
    """The L=100, chi=2048 U(1) Heisenberg DMRG from build_random_symmetric_mps(L, 4, seed)
    (dmrg.py:125-305, 1415-1523): per-sweep energies, truncation errors, seconds and phases.
    compute: the arithmetic of the block GEMMs (matvecs and environments): "f64"
    FP64 MFMA, "f32" FP32 MFMA, "split_f32" compensated FP32; Lanczos
    recurrences, re-orthogonalisation, tridiagonal solve and SVD stay FP64."""
    import torch

    from tenax_amd import DMRGConfig, build_mpo_heisenberg, build_random_symmetric_mps, dmrg

    cfg = DMRGConfig(max_bond_dim=chi, num_sweeps=sweeps, lanczos_max_iter=50, lanczos_reorth=reorth,
                     compute=compute)
    mpo_tn = build_mpo_heisenberg(L)
    mps_tn = build_random_symmetric_mps(L, 4, seed=seed)
    pd.TIMERS.clear()
    pd.TIMERS["enabled"] = True
    torch.cuda.synchronize()
    t0 = time.time()
    res = dmrg(mpo_tn, mps_tn, cfg)
    torch.cuda.synchronize()
    total = time.time() - t0
    phases = {k: v for k, v in pd.TIMERS.items() if k != "enabled"}
    pd.TIMERS["enabled"] = False
    n_upd = 2 * (L - 1)
    errs = res.truncation_errors
    mps = [res.mps.get_tensor(i) for i in range(L)]
    mpo = [mpo_tn.get_tensor(i) for i in range(L)]
    bond_dims = [int(len(mps[i].indices[-1].charges)) for i in range(L - 1)]
    arith = {"f64": "FP64 MFMA", "f32": "FP32 MFMA", "split_f32": "split-FP32 MFMA"}[compute]
    out = {"workload": f"U(1) Heisenberg chain L={L}, Jz=Jxy=1, MPO bond 5, build_random_symmetric_mps(L, 4, "
                       f"seed={seed}), max_bond_dim={chi}, {sweeps} sweeps, Lanczos 50 steps"
                       f"{' with FP64 re-orthogonalisation' if reorth else ''}, {arith} block GEMMs, FP64 SVD",
           "seconds": total, "sweep_seconds": list(pd.LAST_RUN["sweep_seconds"]),
           "energies_per_sweep": list(res.energies_per_sweep), "energy": res.energy,
           "energy_per_site": res.energy / L, "converged": res.converged,
           "max_truncation_error_per_sweep": [max(errs[k * n_upd:(k + 1) * n_upd]) for k in
                                              range(len(res.energies_per_sweep))],
           "max_bond_dim_reached": max(bond_dims), "phases_s": phases}
    return out, mps, mpo, res.mps, mpo_tn


def c3_centre_parity(pd, mps, mpo, reorth):
    """Centre-bond two-site Lanczos of an MPS on the device vs the FP64 oracle's
    Lanczos (stage-wise per-block products) on the same L, W, theta and R."""
    import torch

    from oracle import dmrg as od
    from tests.helpers import to_oracle

    c, left, right = centre_envs(pd, mps, mpo)
    theta, _, _ = pd._theta_padded(mps[c], mps[c + 1])
    prob = pd.TwoSiteProblem(theta, left, mpo[c], mpo[c + 1], right)
    oth = to_oracle(prob.extract(prob.present_after(1)))
    ops = [to_oracle(t) for t in (left, mpo[c], mpo[c + 1], right)]
    torch.cuda.synchronize()
    t0 = time.time()
    e, n = prob.lanczos(50, 1e-12, reorth, 0)
    torch.cuda.synchronize()
    t_gpu = time.time() - t0
    t0 = time.time()
    e_ref, _ = od.lanczos(od.FactoredMatvec(*ops), oth, 50, 1e-12, reorth)
    t_cpu = time.time() - t0
    st = prob.plan.stats()
    return {"bond": [c, c + 1], "theta_elems": int(prob.total), "theta_blocks": int(len(prob.keys)),
            "flops_per_matvec": st["flops"], "energy": e, "oracle_energy": e_ref, "abs_dE": abs(e - e_ref),
            "lanczos_steps": n, "gpu_s": t_gpu, "oracle_s": t_cpu, "reorth": reorth}, prob


def run_c1_parity():
    """C1 (L=20, chi=64, 10 sweeps, seed 0) on the GPU vs the FP64 oracle's golden trajectory."""
    from tenax_amd import DMRGConfig, build_mpo_heisenberg, build_random_symmetric_mps, dmrg

    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "c1_oracle.json")))
    t0 = time.time()
    res = dmrg(build_mpo_heisenberg(20), build_random_symmetric_mps(20, 4, seed=0),
               DMRGConfig(max_bond_dim=64, num_sweeps=10, lanczos_max_iter=50, lanczos_reorth=True))
    dt = time.time() - t0
    ref = gold["reorth"]["energies_per_sweep"]
    n = min(len(ref), len(res.energies_per_sweep))
    dE = max(abs(a - b) for a, b in zip(res.energies_per_sweep[:n], ref[:n]))
    return {"workload": "L=20 chi=64 10 sweeps seed 0, FP64 MFMA, Lanczos reorth",
            "energy": res.energy, "oracle_energy": gold["reorth"]["energy"],
            "max_abs_dE_per_sweep": dE, "dE_final": abs(res.energy - gold["reorth"]["energy"]),
            "ed_anchor": gold["ed_anchor_2sz0"], "sweeps": len(res.energies_per_sweep), "seconds": dt}


def run_hotrg(args):
    """64-temperature Z2 HOTRG at chi=48, 30 steps; free energy vs Onsager."""
    import torch

    H = importlib.import_module("tenax_amd.algorithms.hotrg")
    from tenax_amd.algorithms.trg import compute_ising_tensor, ising_free_energy_exact
    from tenax_amd.contraction import contractor

    n, chi, steps = args.temps, args.hotrg_chi, args.hotrg_steps
    temps = [2.0 + 0.6 * k / max(1, n - 1) for k in range(n)]
    cfg = H.HOTRGConfig(max_bond_dim=chi, num_steps=steps)
    H.hotrg_batched([compute_ising_tensor(1 / temps[0], symmetric=True)], H.HOTRGConfig(chi, 8))  # warm-up
    Ts = [compute_ising_tensor(1 / T, symmetric=True) for T in temps]
    contractor.WORK.update(enabled=True, flops=0.0, launches=0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lnz = H.hotrg_batched(Ts, cfg)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    flops = contractor.WORK["flops"]
    contractor.WORK["enabled"] = False
    f = [-l * T for l, T in zip(lnz, temps)]
    ex = [ising_free_energy_exact(1 / T) for T in temps]
    rel = [abs(a - b) / abs(b) for a, b in zip(f, ex)]
    log(f"hotrg {n} temps chi={chi} {steps} steps: {dt:.2f}s, {flops:.3e} flops, max rel err vs Onsager "
        f"{max(rel):.2e}")
    # phase split on a second, timed-by-phase run of a few instances
    H.TIMERS.clear()
    H.TIMERS["enabled"] = True
    H.hotrg_batched(Ts[: min(8, n)], cfg)
    phases = {k: v for k, v in H.TIMERS.items() if k != "enabled"}
    H.TIMERS["enabled"] = False
    result = {
        "metric": "batched Z2 HOTRG GFLOP/s (block contractions), 64 Ising temperatures", "value": flops / dt / 1e9,
        "unit": "GFLOP/s", "n_gpus": 1, "steps": steps, "warmup": 1, "ms_per_step": dt / steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"2D Ising Z2 HOTRG, {n} temperatures in [2.0, 2.6], chi={chi}, {steps} steps, "
                               "alternating directions, all instances in lockstep", "temperatures": n, "chi": chi},
        "seconds": dt, "contraction_flops": flops, "launches": contractor.WORK["launches"],
        "phases_s_first8": phases,
        "free_energy": {"max_rel_err_vs_onsager": max(rel), "per_T": [[T, a, b] for T, a, b in zip(temps, f, ex)]},
    }
    if not args.no_cpu:
        result["cpu_baseline"] = hotrg_cpu_baseline(chi)
    return result


def hotrg_cpu_baseline(chi):
    """One chi-sized horizontal HOTRG step of the oracle (per-block NumPy
    contractions, LAPACK SVD) vs the same step on the GPU, same input."""
    import numpy as np
    import torch

    from oracle import core as oc
    from oracle import hotrg as oh
    H = importlib.import_module("tenax_amd.algorithms.hotrg")
    from tenax_amd.algorithms.trg import compute_ising_tensor
    from tenax_amd.contraction import contractor

    T = compute_ising_tensor(1 / 2.27, symmetric=True)
    k = 0
    while min(int(ix.dim) for ix in T.indices) < chi:
        T, _ = (H._hotrg_step_horizontal if k % 2 == 0 else H._hotrg_step_vertical)(T, chi)
        k += 1
    if k % 2:
        T, _ = H._hotrg_step_vertical(T, chi)
    host = T.data.cpu().numpy()
    legs = tuple(oc.Leg(oc.Zn(2), ix.charges, int(ix.flow), ix.label) for ix in T.indices)
    blocks = {tuple(int(q) for q in key): host[int(o):int(o) + int(np.prod(sh))].reshape(tuple(int(x) for x in sh))
              for key, sh, o in zip(T.keys, T.shapes, T.offsets)}
    To = oc.SymT(blocks, legs, validate=False)
    contractor.WORK.update(enabled=True, flops=0.0, launches=0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, ln_gpu = H._hotrg_step_horizontal(T, chi)
    torch.cuda.synchronize()
    gpu_s = time.perf_counter() - t0
    flops = contractor.WORK["flops"]
    contractor.WORK["enabled"] = False
    t0 = time.perf_counter()
    _, ln_cpu = oh.step_horizontal(To, chi, order="chi5")
    cpu_s = time.perf_counter() - t0
    log(f"hotrg cpu step {cpu_s:.2f}s gpu step {gpu_s:.3f}s |dln| {abs(ln_cpu - ln_gpu):.1e}")
    return {"value": flops / cpu_s / 1e9, "unit": "GFLOP/s", "cores": os.cpu_count(), "kind": "port",
            "sample": f"one horizontal HOTRG step at chi={chi} (T=2.27) by the oracle (per-block NumPy "
                      "contractions + LAPACK SVD, OpenBLAS threads = cores); value = GPU planner flops / CPU time",
            "seconds": cpu_s, "gpu_seconds_same_step": gpu_s, "abs_diff_log_norm": abs(ln_cpu - ln_gpu)}


def run_ctm(args):
    import numpy as np
    import torch

    from tenax_amd.contraction import contractor
    from tenax_amd.core.tensor import PRNGKey, normal

    C = importlib.import_module("tenax_amd.algorithms.ctm")
    D, d, chi = 4, 2, args.ctm_chi
    A = normal(PRNGKey(3), (D, D, D, D, d))
    A = A / np.linalg.norm(A)
    H = C.heisenberg_gate()
    res = {}
    for name, sym in (("dense", False), ("block", True)):
        At = C.ipeps_site_tensor(A, sym)
        a = C._build_double_layer_tensor(At)
        env = C.initialize_ctm_tensor_env(At, chi)
        with torch.no_grad():
            for _ in range(2):
                env = C._ctm_tensor_sweep(env, a, chi, True, "qr")
            contractor.WORK.update(enabled=True, flops=0.0, launches=0)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            n = args.ctm_sweeps
            for _ in range(n):
                env = C._ctm_tensor_sweep(env, a, chi, True, "qr")
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t0) / n
        res[name] = {"sweep_ms": dt * 1e3}
        if sym:
            res[name]["sweep_flops"] = contractor.WORK["flops"] / n
            res[name]["launches_per_sweep"] = contractor.WORK["launches"] / n
        contractor.WORK["enabled"] = False
        log(f"ctm {name}: {dt * 1e3:.1f} ms per sweep")
    flops = res["block"]["sweep_flops"]
    # AD forward (100 sweeps + gauge fix) and implicit backward on the block path
    cfg = C.CTMConfig(chi=chi, max_iter=100, conv_tol=1e-10, projector_method="qr")
    At = C.ipeps_site_tensor(A, True)
    leaf = C._leaves([At])[0].detach().clone().requires_grad_(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e = C.energy_loss(C._with_leaves([At], [leaf])[0], H, cfg)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    log(f"ctm AD forward {t_fwd:.2f}s E={float(e.detach()):.10f} iterations "
        f"{C.LAST_BACKWARD.get('forward_iterations')}")
    # the reference CTM does not reach a fixed point here (its corner legs swap
    # roles between moves), so GMRES cannot converge: time one restart cycle
    C.GMRES_OPTIONS.update(maxiter=args.gmres_cycles, progress=True)
    t0 = time.perf_counter()
    (g,) = torch.autograd.grad(e, [leaf])
    torch.cuda.synchronize()
    t_bwd = time.perf_counter() - t0
    log(f"ctm AD backward {t_bwd:.2f}s gmres products {C.LAST_BACKWARD.get('gmres_products')} "
        f"residual {C.LAST_BACKWARD.get('relative_residual'):.2e}")
    result = {
        "metric": "iPEPS CTMRG sweep GFLOP/s (block path), D=4 chi=64", "value": flops / (res["block"]["sweep_ms"] * 1e-3)
        / 1e9, "unit": "GFLOP/s", "n_gpus": 1, "steps": args.ctm_sweeps, "warmup": 2,
        "ms_per_step": res["block"]["sweep_ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"1-site iPEPS, Heisenberg, D=4, d=2, CTMRG chi={chi}, QR projectors, A ~ N(0,1) "
                               "seed 3 normalised", "chi": chi, "D": D},
        "sweep": res, "ad": {"forward_s": t_fwd, "backward_s": t_bwd, "energy": float(e.detach()),
                             "gmres_cycles": args.gmres_cycles,
                             "backward_s_per_gmres_product": t_bwd / max(1, C.LAST_BACKWARD.get("gmres_products", 1)),
                             "forward_iterations": C.LAST_BACKWARD.get("forward_iterations"),
                             "gmres_products": C.LAST_BACKWARD.get("gmres_products"),
                             "relative_residual": C.LAST_BACKWARD.get("relative_residual"),
                             "grad_norm": float(torch.linalg.vector_norm(g))},
    }
    if not args.no_cpu:
        from oracle import ctm as occ

        a_np = occ.double_layer(A)
        env_np = occ.init_env(D * D, chi)
        env_np = occ.sweep(env_np, a_np, chi, True, "qr")
        t0 = time.perf_counter()
        env_np = occ.sweep(env_np, a_np, chi, True, "qr")
        cpu = time.perf_counter() - t0
        result["cpu_baseline"] = {"value": flops / cpu / 1e9, "unit": "GFLOP/s", "cores": os.cpu_count(),
                                  "kind": "port", "seconds": cpu,
                                  "sample": f"one CTM sweep at chi={chi} by the oracle (NumPy FP64 einsum / LAPACK, "
                                            "OpenBLAS threads = cores); value = block-path flops / CPU time"}
        log(f"ctm cpu sweep {cpu:.2f}s")
    return result


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--L", type=int, default=100)
    ap.add_argument("--chi", type=int, default=2048)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-parity", action="store_true")
    ap.add_argument("--perblock-steps", type=int, default=2)
    ap.add_argument("--no-sweep", dest="sweep", action="store_false",
                    help="skip the L=100 chi=2048 DMRG run from a random MPS (run by default)")
    ap.add_argument("--c3-sweeps", type=int, default=6)
    ap.add_argument("--perblock-sweep", action="store_true",
                    help="also time one full sweep from the C3 state on both paths")
    ap.add_argument("--no-reorth", action="store_true", help="C3 with the reference's plain Lanczos")
    ap.add_argument("--workload", choices=("dmrg", "hotrg", "ctm"), default="dmrg")
    ap.add_argument("--no-c4c5", dest="c4c5", action="store_false",
                    help="leave the chi=1024 contraction (c2), HOTRG (c5) and CTMRG (c4) summary keys out of the "
                         "DMRG line")
    ap.add_argument("--matvec-only", action="store_true", help="fused matvec and stage timing only (profiling)")
    ap.add_argument("--calibrate-copy", action="store_true", help="run a 512 MiB device copy (counter calibration)")
    ap.add_argument("--pin-tiling", action="append", default=[], metavar="STAGE:SHAPE,KCHUNK,XCD",
                    help="fix a matvec GEMM stage's tiling (counter runs, whose timing would mislead the autotuner)")
    ap.add_argument("--ctm-chi", type=int, default=64)
    ap.add_argument("--ctm-sweeps", type=int, default=5)
    ap.add_argument("--gmres-cycles", type=int, default=1)
    ap.add_argument("--temps", type=int, default=64)
    ap.add_argument("--hotrg-chi", type=int, default=48)
    ap.add_argument("--hotrg-steps", type=int, default=30)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "hotrg":
        print(json.dumps(run_hotrg(args)), flush=True)
        return
    if args.workload == "ctm":
        print(json.dumps(run_ctm(args)), flush=True)
        return
    import torch

    from tenax_amd import _lib

    pd, prob, (L, Wl, Wr, R, theta), mps, mpo = build_problem(args.L, args.chi, args.seed)
    for pin in args.pin_tiling:
        st, cfg = pin.split(":")
        prob.plan.set_tiling(int(st), *(int(v) for v in cfg.split(",")))
    stats = prob.plan.stats()
    flops = stats["flops"]
    log(f"matvec plan: flops={flops:.3e} pairs={stats['n_pairs']} tasks={stats['n_tasks']}")

    # fused FP64 matvec: the timed steps are replayed from one captured graph
    # of `steps` matvecs, the way the sweep's Lanczos solves run them (each
    # solve is one captured graph); the same matvecs launched eagerly are
    # timed as well
    for _ in range(args.warmup):
        prob.matvec(prob.x, 0)
    ms_eager = time_events(lambda: prob.matvec(prob.x, 0), args.steps)
    mv_graph, y_graph = capture_matvecs(prob, args.steps)
    mv_graph.replay()
    ms = time_events(mv_graph.replay, 1) / args.steps
    y_eager = prob.matvec(prob.x, 0)
    torch.cuda.synchronize()
    graph_equal = bool(torch.equal(y_graph, y_eager))
    del mv_graph, y_graph
    gflops = flops / (ms * 1e-3) / 1e9
    log(f"fused f64 matvec {ms:.3f} ms  {gflops:.1f} GFLOP/s (captured; eager launches {ms_eager:.3f} ms, "
        f"outputs equal {graph_equal})")

    # per-stage kernel times -> roofline of the dominant kernel
    st_ms, st_fl = stage_times(prob, max(10, args.steps), ms)
    dom = max((1, 3), key=lambda s: st_ms[f"{s}_gemm"])
    ach = st_fl[dom] / (st_ms[f"{dom}_gemm"] * 1e-3) / 1e12
    log(f"stages ms {st_ms}  flops {st_fl}")

    if args.calibrate_copy:
        # a plain device copy of known size (512 MiB read + 512 MiB written),
        # the reference launch for scaling FETCH_SIZE / WRITE_SIZE counters
        src = torch.ones(64 << 20, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.matvec_only:
        print(json.dumps({"matvec_ms": ms, "gflops": gflops, "stage_ms": st_ms, "stage_flops": st_fl,
                          "kernels": {st: stage_kernel(prob, st)[:2] for st in (1, 3)}}), flush=True)
        return

    # reduced precision variants
    y64 = prob.matvec(prob.x, 0)
    extra = {}
    for mode in ("f32", "split_f32"):
        c = _lib.COMPUTE[mode]
        prob.matvec(prob.x, c)
        m2 = time_events(lambda: prob.matvec(prob.x, c), max(3, args.steps // 2))
        y = prob.matvec(prob.x, c)
        err = float((y - y64).abs().max() / y64.abs().max())
        extra[mode] = {"gflops": flops / (m2 * 1e-3) / 1e9, "ms": m2, "max_rel_err_vs_f64": err}
        log(f"{mode}: {m2:.3f} ms rel err {err:.2e}")

    # plain per-block GPU path (one launch per block product)
    prob.matvec(prob.x, 0, per_block=True)
    ms_pb = time_events(lambda: prob.matvec(prob.x, 0, per_block=True), args.perblock_steps)
    log(f"per-block matvec {ms_pb:.3f} ms")

    # one full device Lanczos solve (50 steps) on the same problem
    x0 = prob.x.clone()
    torch.cuda.synchronize()
    t0 = time.time()
    e, n = prob.lanczos(50, 1e-12, False, 0)
    torch.cuda.synchronize()
    lz_s = time.time() - t0
    prob.x = x0
    log(f"lanczos 50 steps {lz_s:.3f}s E={e:.12f} n={n}")

    # the same solve recorded once as a hipGraph, then replayed
    torch.cuda.synchronize()
    t0 = time.time()
    graph = prob.lanczos_graph(50, 1e-12, False, 0)
    torch.cuda.synchronize()
    cap_s = time.time() - t0
    xs = prob.x.clone()
    graph.launch()
    graph.result()
    prob.x.copy_(xs)
    torch.cuda.synchronize()
    t0 = time.time()
    graph.launch()
    eg, ng = graph.result()
    lzg_s = time.time() - t0
    graph.close()
    prob.x.copy_(xs)
    log(f"lanczos graph: capture+instantiate {cap_s:.3f}s, replay {lzg_s:.3f}s E={eg:.12f} (eager E diff {eg - e:.1e})")

    kname, kwgs, _ = stage_kernel(prob, dom)
    traffic = committed_traffic(kname, kwgs) if (args.L == 100 and args.chi == 2048) else None
    tiling = {f"stage{st}": stage_kernel(prob, st)[2] for st in (1, 3)}
    result = {
        "metric": METRIC,
        "value": gflops,
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"fused two-site matvec, U(1) Heisenberg chain L={args.L}, chi={args.chi}, "
                               f"MPO bond 5, centre bond ({args.L // 2 - 1},{args.L // 2}), converged-profile "
                               f"bond charges, seed {args.seed}",
                   "L": args.L, "chi": args.chi, "theta_blocks": int(len(prob.keys)),
                   "theta_elems": int(prob.total), "flops_per_matvec": flops,
                   "block_pairs": stats["n_pairs"], "gemm_tiles": stats["n_tasks"]},
        "tiling": tiling,
        "roofline": {"bound": "mfma", "kernel": f"{kname} x {kwgs} (matvec stage {dom}, FP64 MFMA)",
                     "achieved": ach, "peak": PEAK_FP64_MFMA_TFLOPS, "unit": "TFLOP/s",
                     "frac": ach / PEAK_FP64_MFMA_TFLOPS, "traffic": traffic,
                     "kernel_avg_us": st_ms[f"{dom}_gemm"] * 1e3,
                     "kernel_timing": f"in place: {max(10, args.steps)} matvecs launched kernel by kernel with a "
                                      f"fence-free HIP event (hipEventDisableSystemFence) before and after every "
                                      f"kernel; each kernel's event share of the timed (captured) matvec time "
                                      f"(stage_ms['events'] raw, stage_ms['isolated'] each kernel back to back "
                                      f"on its own)",
                     "traffic_source": "profiles/dmrg_matvec_chi2048_traffic.md (rocprofv3 --pmc FETCH_SIZE x2.0 + "
                                       "WRITE_SIZE x1.0, copy-calibrated)" if traffic else None,
                     "stage_ms": st_ms, "stage_flops": st_fl},
        "matvec_timing": {"captured_ms": ms, "eager_ms": ms_eager, "eager_gflops": flops / (ms_eager * 1e-3) / 1e9,
                          "captured_equals_eager": graph_equal,
                          "method": f"{args.steps} matvecs replayed from one captured graph (HIP events around the "
                                    f"replay, synchronised before and after); eager: the same matvecs launched "
                                    f"one by one"},
        "per_block_gpu": {"ms": ms_pb, "gflops": flops / (ms_pb * 1e-3) / 1e9, "speedup_fused": ms_pb / ms,
                          "speedup_fused_eager": ms_pb / ms_eager},
        "reduced_precision": extra,
        "lanczos_50_steps_s": lz_s,
        "lanczos_50_steps_graph": {"replay_s": lzg_s, "capture_instantiate_s": cap_s,
                                   "energy_equal_to_eager": eg == e},
        "sweep_s": None,
    }
    if args.sweep:
        c3, c3_mps, c3_mpo, c3_mps_tn, c3_mpo_tn = run_c3(pd, args.L, args.chi, args.c3_sweeps, args.seed,
                                                          not args.no_reorth)
        result["c3"] = c3
        result["sweep_s"] = c3["sweep_seconds"][-1]
        log(f"C3 {c3['seconds']:.1f}s sweeps {c3['sweep_seconds']} E {c3['energies_per_sweep']} "
            f"phases {c3['phases_s']}")
        par, cprob = c3_centre_parity(pd, c3_mps, c3_mpo, not args.no_reorth)
        # the headline matvec timed on the C3 state's own centre bond as well
        for _ in range(args.warmup):
            cprob.matvec(cprob.x, 0)
        cms = time_events(lambda: cprob.matvec(cprob.x, 0), args.steps)
        par["matvec_ms"] = cms
        par["matvec_gflops"] = par["flops_per_matvec"] / (cms * 1e-3) / 1e9
        result["c3"]["centre_bond"] = par
        log(f"C3 centre bond: {par}")
        del cprob
        result["c3"]["centre_bond_reduced_precision"] = c3_reduced_precision(pd, c3_mps, c3_mpo, not args.no_reorth)
        log(f"C3 centre bond reduced precision: {result['c3']['centre_bond_reduced_precision']}")
        # the whole C3 run again with the block GEMMs in FP32 and in split FP32
        runs = {}
        for mode in ("f32", "split_f32"):
            r, *_ = run_c3(pd, args.L, args.chi, args.c3_sweeps, args.seed, not args.no_reorth, compute=mode)
            runs[mode] = {"seconds": r["seconds"], "energies_per_sweep": r["energies_per_sweep"],
                          "energy": r["energy"], "abs_dE_vs_f64_run": abs(r["energy"] - c3["energy"]),
                          "phases_s": r["phases_s"]}
            log(f"C3 {mode}: {runs[mode]}")
        result["c3"]["reduced_precision_runs"] = runs
        cmp = compare_updates(pd, c3_mps, c3_mpo, args.chi, not args.no_reorth)
        result["two_site_updates_centre_s"] = cmp
        log(f"C3 centre updates fused vs per-block: {cmp}")
        # one full sweep at chi = 2048 everywhere (the C3 run's last left-to-right
        # half still had chi = 1024 bonds), and the same sweep on the per-block
        # path (about 3 minutes; a 20-bond centre segment by default)
        if args.perblock_sweep:
            sc = sweep_comparison(pd, c3_mps_tn, c3_mpo_tn, args.chi, not args.no_reorth)
        else:
            sc = sweep_comparison(pd, c3_mps_tn, c3_mpo_tn, args.chi, not args.no_reorth, fused_only=True)
            sc["segment"] = segment_comparison(pd, c3_mps, c3_mpo, args.chi, not args.no_reorth)
        result["sweep_vs_per_block"] = sc
        result["sweep_s"] = sc["fused"]["seconds"]
    if not args.no_cpu:
        cb = cpu_baseline(prob)
        result["cpu_baseline"] = {
            "value": flops / cb["seconds"] / 1e9, "unit": "GFLOP/s", "cores": cb["threads"], "kind": "port",
            "sample": "1 matvec of the headline centre-bond problem by the oracle's per-block-combination loop "
                      "(the reference's loop structure; NumPy FP64, OpenBLAS), 1 warm-up + 3 trials, median; "
                      "value = planner flops / CPU time",
            "single_thread_value": flops / cb["single_thread_seconds_est"] / 1e9,
            "single_thread_sample": "every 8th theta block by size (1 warm-up + 3 trials at 1 and at all "
                                    "threads); full-matvec time scaled by the sample's 1-thread/all-thread ratio",
            **{k: cb[k] for k in ("seconds", "seconds_trials", "sample_seconds_all_threads",
                                  "sample_seconds_1_thread", "single_thread_seconds_est", "max_rel_diff_vs_gpu",
                                  "host")}}
        log(f"cpu baseline {cb['seconds']:.2f}s ({cb['threads']} threads), 1 thread est "
            f"{cb['single_thread_seconds_est']:.1f}s, rel diff {cb['max_rel_diff_vs_gpu']:.2e}")
    if not args.no_parity:
        result["parity"] = run_c1_parity()
        log(f"parity {result['parity']}")
    if args.c4c5:
        result["c2"] = c2_summary(args)
        result["c5"] = c5_summary(args)
        result["c4"] = c4_summary(args)
    print(json.dumps(result), flush=True)


def c2_summary(args, reps=200):
    """The single block-sparse contraction at chi=1024 (rank-4 U(1) tensors,
    21-sector Gaussian multiplicities, N(0,1) entries from seed 1, contracted
    over r): structure, grouped-GEMM time per launch (HIP events over
    repeated launches of the cached plan) and GFLOP/s in FP64, FP32 and
    split FP32, with the reduced-precision errors against the FP64 result."""
    import torch

    from tenax_amd import SymmetricTensor, contract
    from tenax_amd.contraction import contractor, plans
    from tenax_amd.core.index import FlowDirection, TensorIndex
    from tenax_amd.core.symmetry import U1Symmetry
    from tenax_amd.workloads import PHYS_CHARGES, gaussian_bond_charges

    U1 = U1Symmetry()
    bond = gaussian_bond_charges(1024)
    IN, OUT = FlowDirection.IN, FlowDirection.OUT
    ia = (TensorIndex(U1, bond, IN, "l"), TensorIndex(U1, PHYS_CHARGES, IN, "p"),
          TensorIndex(U1, PHYS_CHARGES, IN, "pp"), TensorIndex(U1, bond, OUT, "r"))
    ib = (TensorIndex(U1, bond, IN, "r"), TensorIndex(U1, PHYS_CHARGES, IN, "q"),
          TensorIndex(U1, PHYS_CHARGES, IN, "qq"), TensorIndex(U1, bond, OUT, "s"))
    A = SymmetricTensor.random_normal(ia, (1, 0))
    B = SymmetricTensor.random_normal(ib, (1, 1))
    C = contract(A, B)
    plan = plans.pairwise_plan(A, B, 0, 1, 1, C.indices, 0, None)[0]
    flops = plan.stats()["flops"]
    out = {"workload": "A[l,p,p',r] . B[r,q,q',s], U(1), chi=1024 (21 sectors), N(0,1) seed 1",
           "a_blocks": int(A.n_blocks), "a_elems": int(A.data.numel()), "out_blocks": int(C.n_blocks),
           "out_elems": int(C.data.numel()), "block_pairs": plan.stats()["n_pairs"], "flops": flops}
    ref = C.data.clone()
    for mode in ("f64", "f32", "split_f32"):
        with contractor.compute_mode(mode):
            y = torch.empty_like(ref)
            contractor._execute_pair(plan, A, B, y)
            ms = time_events(lambda: contractor._execute_pair(plan, A, B, y), reps)
        out[mode] = {"us": ms * 1e3, "gflops": flops / (ms * 1e-3) / 1e9,
                     "max_rel_err_vs_f64": float((y - ref).abs().max() / ref.abs().max())}
    out["f64"]["frac_of_fp64_mfma_peak"] = out["f64"]["gflops"] / 1e3 / PEAK_FP64_MFMA_TFLOPS
    log(f"c2 chi=1024: {out}")
    return out


def c3_reduced_precision(pd, mps, mpo, reorth):
    """The C3 centre-bond two-site Lanczos with the matvec GEMMs in FP32 MFMA
    and in split FP32 (FP64 Lanczos recurrences, re-orthogonalisation and
    tridiagonal solve), energies against the FP64 solve on the same input."""
    import torch

    c, left, right = centre_envs(pd, mps, mpo)
    theta, _, _ = pd._theta_padded(mps[c], mps[c + 1])
    out = {}
    x0 = None
    for mode in ("f64", "f32", "split_f32"):
        prob = pd.TwoSiteProblem(theta, left, mpo[c], mpo[c + 1], right)
        if x0 is None:
            x0 = prob.x.clone()
        prob.x.copy_(x0)
        torch.cuda.synchronize()
        t0 = time.time()
        e, n = prob.lanczos(50, 1e-12, reorth, _lib_compute(mode))
        torch.cuda.synchronize()
        out[mode] = {"energy": e, "steps": n, "seconds": time.time() - t0}
    for mode in ("f32", "split_f32"):
        out[mode]["abs_dE_vs_f64"] = abs(out[mode]["energy"] - out["f64"]["energy"])
    return out


def _lib_compute(mode):
    from tenax_amd import _lib

    return _lib.COMPUTE[mode]


def c5_summary(args):
    """The batched HOTRG workload (64 Z2 Ising temperatures, chi=48, 30 steps)
    as a key of the default line: time, block-contraction GFLOP/s, phase split
    and the worst free-energy error against Onsager."""
    r = run_hotrg(argparse.Namespace(**{**vars(args), "no_cpu": True}))
    out = {k: r[k] for k in ("metric", "value", "unit", "ms_per_step", "config", "seconds", "contraction_flops",
                             "launches", "phases_s_first8")}
    out["max_rel_err_vs_onsager"] = r["free_energy"]["max_rel_err_vs_onsager"]
    return out


def c4_summary(args):
    """The iPEPS CTMRG workload (D=4, chi=64, QR projectors) as a key of the
    default line: sweep time on the block and dense paths, AD forward and one
    GMRES cycle of the implicit backward."""
    r = run_ctm(argparse.Namespace(**{**vars(args), "no_cpu": True}))
    return {k: r[k] for k in ("metric", "value", "unit", "ms_per_step", "config", "sweep", "ad")}


if __name__ == "__main__":
    main()

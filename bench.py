"""Benchmark of the block-sparse Contract hot path on one MI355X.

Workload: the two-site effective-Hamiltonian matvec psi' = L.psi.W.W.R on the
U(1)xU(1) (N, 2Sz) Hubbard-like tensors, D = 4096 over 85 sectors, Dmpo = 6,
complex128, uniform(-1,1) entries from a fixed seed (synthetic data).  A
step is one matvec: the four grouped contractions on cached device plans,
inputs resident in HBM.  Throughput = algorithmic flops (sum over tasks of
8*m*k*n, tensor_op_cost.h:384-406) / time.

With ``--autotune``, before each mode is timed the autotuner (qlt_autotune)
chooses tile shapes per size bucket and k-slicing targets on the chain
itself, outside the timed region, and the plans are rebuilt under its
result; by default the built-in tile rules are timed (on this workload the
tuned and built-in chains measure within run-to-run spread of each other).

The JSON line reports the split-BF16 path (``value``) next to the native
FP64 path; the roofline of the dominant kernel (the library's kernel probe
records HIP events around that one kernel launch, on its launch stream,
in every timed step); the error of the step's output against the FP64 CPU
oracle; and the CPU baseline: the oracle (the reference algorithm restated
in C, OpenBLAS zgemm per task) on a bounded sample.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from tensortoolkit_amd import capi, workloads as W  # noqa: E402
from tensortoolkit_amd import autotune as TUNE  # noqa: E402
from tensortoolkit_amd.blocksparse import ContractionPlan  # noqa: E402

PEAK_FP64_TFLOPS = 78.6
PEAK_BF16_TFLOPS = 2500.0
PEAK_INT8_TOPS = 5000.0
PEAK_HBM_GBS = 8000.0


class Hip:
    """hipEvent timing through libamdhip64 (events recorded on the launch stream)."""

    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")
        self.h.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.h.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.h.hipEventSynchronize.argtypes = [C.c_void_p]
        self.h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.h.hipEventDestroy.argtypes = [C.c_void_p]

    def event(self):
        e = C.c_void_p()
        assert self.h.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.h.hipEventRecord(e, stream) == 0

    def elapsed_ms(self, e0, e1) -> float:
        assert self.h.hipEventSynchronize(e1) == 0
        ms = C.c_float()
        assert self.h.hipEventElapsedTime(C.byref(ms), e0, e1) == 0
        return float(ms.value)


class MatvecRunner:
    """Device buffers and cached plans of one matvec chain."""

    def __init__(self, mv: W.Matvec, stream):
        self.mv, self.stream, self.dtype = mv, stream, mv.dtype
        self.esz = 16 if mv.dtype == "c128" else 8
        self.bufs = {name: capi.DeviceBuffer.from_numpy(getattr(mv, name).data) for name in ("L", "psi", "W", "R")}
        t0 = time.perf_counter()
        self.plans, self.outs, self.out_tensors, self.operands_t = [], [], [], []
        x = mv.psi
        for step, axes in enumerate(W.MATVEC_AXES):
            op = (mv.L, mv.W, mv.W, mv.R)[step]
            first, second = (op, x) if step == 0 else (x, op)
            plan = ContractionPlan(first, second, list(axes[0]), list(axes[1]), stream=stream)
            out = plan.output_tensor(self.dtype)
            self.plans.append(plan)
            self.operands_t.append((first, second, list(axes[0]), list(axes[1])))
            self.out_tensors.append(out)
            self.outs.append(capi.DeviceBuffer(max(1, out.raw_size) * self.esz))
            x = out
        capi.check(capi.lib().qlt_stream_sync(stream))
        self.plan_seconds = time.perf_counter() - t0
        self.flops = [p.flops(self.dtype) for p in self.plans]

    def tune_items(self):
        """The chain as an autotuning workload (same operands and outputs)."""
        items = []
        for step, (first, second, a_ax, b_ax) in enumerate(self.operands_t):
            a, b = self.operands(step)
            items.append((first, second, a_ax, b_ax, a, b, self.outs[step].ptr))
        return items

    def rebuild_plans(self):
        """Rebuild the plans under the library tuning now in effect."""
        for p in self.plans:
            p.close()
        self.plans = [ContractionPlan(f, s, a_ax, b_ax, stream=self.stream) for f, s, a_ax, b_ax in self.operands_t]
        capi.check(capi.lib().qlt_stream_sync(self.stream))

    def operands(self, step):
        names = ("L", "W", "W", "R")
        if step == 0:
            return self.bufs["L"].ptr, self.bufs["psi"].ptr
        return self.outs[step - 1].ptr, self.bufs[names[step]].ptr

    def launch(self, step, mode):
        a, b = self.operands(step)
        self.plans[step].execute(self.dtype, mode, a, b, self.outs[step].ptr, stream=self.stream)

    def result(self) -> np.ndarray:
        out = self.out_tensors[-1]
        return self.outs[-1].to_numpy(out.np_dtype, out.raw_size)


def settle(runner: MatvecRunner, mode: int, seconds: float):
    """Runs the matvec untimed, in batches of 10, for `seconds` of wall time,
    so that the GPU clocks have reached their loaded state before warmup (after
    a pause they take about 20 ms of sustained work to come back up)."""
    lib = capi.lib()
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        for _ in range(10):
            for s in range(4):
                runner.launch(s, mode)
        capi.check(lib.qlt_stream_sync(runner.stream))


def time_mode(runner: MatvecRunner, hip: Hip, mode: int, steps: int, warmup: int, probe_step: int,
              settle_s: float = 0.0):
    """Times `steps` matvecs between one event pair on the launch stream:
    the timed region holds the launches and nothing else.  A second pass of
    `steps` matvecs after it records an event pair around every contraction
    (the per-contraction breakdown) and the library's kernel probe around the
    dominant kernel of contraction `probe_step`; those events add gaps
    between kernels, so that pass is not the step time."""
    lib = capi.lib()
    for p in runner.plans:
        p.prepare(runner.dtype, mode)
    settle(runner, mode, settle_s)
    for _ in range(warmup):
        for s in range(4):
            runner.launch(s, mode)
    capi.check(lib.qlt_stream_sync(runner.stream))
    t_start, t_end = hip.event(), hip.event()
    capi.check(lib.qlt_device_sync())
    wall0 = time.perf_counter()
    hip.record(t_start, runner.stream)
    for _ in range(steps):
        for s in range(4):
            runner.launch(s, mode)
    hip.record(t_end, runner.stream)
    capi.check(lib.qlt_stream_sync(runner.stream))
    capi.check(lib.qlt_device_sync())
    wall = time.perf_counter() - wall0
    total_ms = hip.elapsed_ms(t_start, t_end)
    # breakdown pass
    ev = [[(hip.event(), hip.event()) for _ in range(4)] for _ in range(steps)]
    kev = [(hip.event(), hip.event()) for _ in range(steps)]
    name = C.create_string_buffer(96)
    kflops, kbytes = C.c_double(), C.c_double()
    for k in range(steps):
        for s in range(4):
            hip.record(ev[k][s][0], runner.stream)
            if s == probe_step:
                capi.check(lib.qlt_kernel_probe_set(*kev[k]))
            runner.launch(s, mode)
            if s == probe_step:
                capi.check(lib.qlt_kernel_probe_last(name, 96, C.byref(kflops), C.byref(kbytes)))
                capi.check(lib.qlt_kernel_probe_set(None, None))
            hip.record(ev[k][s][1], runner.stream)
    capi.check(lib.qlt_stream_sync(runner.stream))
    per_launch = np.array([[hip.elapsed_ms(*ev[k][s]) for s in range(4)] for k in range(steps)])
    kernel = {"name": name.value.decode(), "flops": kflops.value, "bytes": kbytes.value,
              "ms": float(np.mean([hip.elapsed_ms(*kev[k]) for k in range(steps)]))}
    return total_ms, wall, per_launch.mean(axis=0), kernel


def time_graph(runner: MatvecRunner, hip: Hip, mode: int, replays: int):
    lib = capi.lib()
    capi.check(lib.qlt_graph_begin(runner.stream))
    for s in range(4):
        runner.launch(s, mode)
    g = C.c_void_p()
    capi.check(lib.qlt_graph_end(runner.stream, C.byref(g)))
    capi.check(lib.qlt_graph_replay(g, runner.stream))
    capi.check(lib.qlt_stream_sync(runner.stream))
    e0, e1 = hip.event(), hip.event()
    hip.record(e0, runner.stream)
    for _ in range(replays):
        capi.check(lib.qlt_graph_replay(g, runner.stream))
    hip.record(e1, runner.stream)
    ms = hip.elapsed_ms(e0, e1)
    capi.check(lib.qlt_graph_destroy(g))
    return ms / replays


def config2_graph_replay(hip: Hip, stream, replays: int, tune: bool) -> dict:
    """The lognormal-degeneracy U(1) FP64 matvec (sector degeneracies
    lognormal(4, 1) clipped to 1..512): the 4-contraction chain captured once
    as a hipGraph and replayed `replays` times, after the autotuner has chosen
    tile shapes and k-slicing on the chain itself (untimed); the eager chain
    under the same tuning beside it."""
    mv2 = W.config2()
    r2 = MatvecRunner(mv2, stream)
    flops = float(sum(r2.flops))
    out = {"workload": mv2.name, "dtype": "f64", "mode": "native", "replays": replays,
           "gflop_per_step": round(flops / 1e9, 4), "tasks": [int(p.info.n_tasks) for p in r2.plans]}
    TUNE.set_tuning(None)
    if tune:
        t0 = time.perf_counter()
        res = TUNE.autotune(r2.tune_items(), r2.dtype, capi.QLT_MODE_NATIVE, reps=3, max_buckets=16, stream=stream)
        out["autotune"] = {"trials": res.trials, "ms_before": round(res.ms_before, 4),
                           "ms_after": round(res.ms_after, 4), "overrides": TUNE.overrides(res.tuning),
                           "seconds": round(time.perf_counter() - t0, 1)}
        r2.rebuild_plans()
    total_ms, _, _, _ = time_mode(r2, hip, capi.QLT_MODE_NATIVE, 20, 3, int(np.argmax(r2.flops)), 0.1)
    graph_ms = time_graph(r2, hip, capi.QLT_MODE_NATIVE, replays)
    out.update(ms_per_step_eager=round(total_ms / 20, 4), gflops_eager=round(flops / (total_ms / 20 * 1e-3) / 1e9, 2),
               ms_per_replay=round(graph_ms, 4), gflops_replay=round(flops / (graph_ms * 1e-3) / 1e9, 2))
    TUNE.set_tuning(None)
    for p in r2.plans:
        p.close()
    return out


# Slice products per algorithmic product in the split-BF16 GEMM: five BF16
# slices per operand, every slice pair (i, j) with i + j <= 4.
SPLIT_PRODUCTS = 15
# Digit products per algorithmic product in the int8 split GEMM: six
# base-255 digits per operand, every digit pair (i, j) with i + j <= 5.
I8_PRODUCTS = 21


def profile_traffic(kernel_name: str, workload: str):
    """HBM bytes per dispatch of `kernel_name` on `workload`, read from the
    committed rocprofv3 PMC summary (profiles/*.pmc.txt: FETCH_SIZE and
    WRITE_SIZE per dispatch, scaled by a device-to-device copy of known
    size), or None when no summary covers this kernel and workload."""
    d = os.path.join(ROOT, "profiles")
    if not os.path.isdir(d):
        return None, None
    for fn in sorted(os.listdir(d)):
        if not fn.endswith(".pmc.txt"):
            continue
        kv = {}
        with open(os.path.join(d, fn)) as f:
            for line in f:
                if ":" in line and not line.startswith(" "):
                    k, v = line.split(":", 1)
                    kv[k.strip()] = v.strip()
        if kv.get("kernel") == kernel_name and kv.get("workload") == workload and "traffic_bytes_per_dispatch" in kv:
            return float(kv["traffic_bytes_per_dispatch"]), f"profiles/{fn}"
    return None, None


def kernel_roofline(kernel: dict, workload: str) -> dict:
    """Roofline of the probed kernel: algorithmic flops (sum of c*m*k*n over
    its tasks, the reference's AddGemmCost) over its event-timed duration.
    A split GEMM runs on BF16 (int8) MFMA, so `achieved` (algorithmic flops
    per second, no slice products counted) is set against the BF16 (int8)
    peak; the rate of the slice products the units actually execute (15 BF16
    or 21 int8 products per algorithmic product) is reported beside it, as
    is the FP64-equivalent fraction of the FP64 peak."""
    secs = kernel["ms"] * 1e-3
    fp64_tf = kernel["flops"] / secs / 1e12
    split = kernel["name"].startswith("k_split_gemm")
    split_i8 = kernel["name"].startswith("k_split_i8")
    rows = kernel["name"].startswith("k_rows")
    traffic, src = profile_traffic(kernel["name"], workload)
    out = {"kernel": kernel["name"], "ms_per_launch": round(kernel["ms"], 5),
           "algorithmic_flops": kernel["flops"], "algorithmic_bytes": kernel["bytes"]}
    if rows:
        gbs = kernel["bytes"] / secs / 1e9
        out.update(bound="hbm", achieved=round(gbs, 1), peak=PEAK_HBM_GBS, unit="GB/s",
                   frac=round(gbs / PEAK_HBM_GBS, 4))
    elif split_i8:
        i8_tops = I8_PRODUCTS * fp64_tf
        out.update(bound="mfma", achieved=round(fp64_tf, 3), peak=PEAK_INT8_TOPS, unit="TOP/s",
                   frac=round(fp64_tf / PEAK_INT8_TOPS, 5),
                   slice_products_per_algorithmic_product=I8_PRODUCTS,
                   slice_product_rate=round(i8_tops, 2), slice_product_frac=round(i8_tops / PEAK_INT8_TOPS, 4),
                   fp64_equivalent_tflops=round(fp64_tf, 3),
                   fp64_equivalent_frac_of_fp64_peak=round(fp64_tf / PEAK_FP64_TFLOPS, 4))
    elif split:
        bf16_tf = SPLIT_PRODUCTS * fp64_tf
        out.update(bound="mfma", achieved=round(fp64_tf, 3), peak=PEAK_BF16_TFLOPS, unit="TFLOP/s",
                   frac=round(fp64_tf / PEAK_BF16_TFLOPS, 5),
                   slice_products_per_algorithmic_product=SPLIT_PRODUCTS,
                   slice_product_rate=round(bf16_tf, 2), slice_product_frac=round(bf16_tf / PEAK_BF16_TFLOPS, 4),
                   fp64_equivalent_tflops=round(fp64_tf, 3),
                   fp64_equivalent_frac_of_fp64_peak=round(fp64_tf / PEAK_FP64_TFLOPS, 4))
    else:
        out.update(bound="mfma", achieved=round(fp64_tf, 3), peak=PEAK_FP64_TFLOPS, unit="TFLOP/s",
                   frac=round(fp64_tf / PEAK_FP64_TFLOPS, 4))
    out["traffic"] = traffic
    if src:
        out["traffic_source"] = src
    return out


def cpu_baseline(mv: W.Matvec, budget_s: float = 10.0):
    """The oracle (reference algorithm restated in C, OpenBLAS per task) on a
    bounded sample: whole matvec steps until about `budget_s` seconds."""
    from oracle import oracle as O
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores) or cores))
    info = O.use_openblas(cores)
    # plans are built once (the reference rebuilds them per call; timed separately)
    t0 = time.perf_counter()
    x = mv.psi
    plans = []
    for step, axes in enumerate(W.MATVEC_AXES):
        op = (mv.L, mv.W, mv.W, mv.R)[step]
        first, second = (op, x) if step == 0 else (x, op)
        p = O.OraclePlan(first, second, list(axes[0]), list(axes[1]))
        out = p.output_tensor(mv.dtype)
        plans.append((p, first is op))
        x = out
    plan_s = time.perf_counter() - t0
    flops_step = sum((8.0 if mv.dtype == "c128" else 2.0) * float((p.tasks()["m"] * p.tasks()["k"] * p.tasks()["n"]).sum())
                     for p, _ in plans)
    done, t_exec, result = 0, 0.0, None
    while True:
        t0 = time.perf_counter()
        x = mv.psi
        for step, (p, _) in enumerate(plans):
            if step > 0:
                p.a.data = x
            out = p.execute()
            x = out
        t_exec += time.perf_counter() - t0
        done += 1
        result = x
        if t_exec >= budget_s or done >= 50:
            break
    return {
        "value": flops_step * done / t_exec / 1e9,
        "unit": "GFLOP/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{done} full matvec step(s), execute only ({t_exec:.1f} s); plan build {plan_s:.2f} s "
                  f"once; OpenBLAS {'zgemm' if mv.dtype == 'c128' else 'dgemm'} per task, "
                  f"{info['threads']} threads ({info['library']})",
        "cpu_model": cpu_model(),
        "blas": info.get("config") or info["library"],
        "threads_note": f"{info['threads']} BLAS threads = the host cores this process may use; the reference's "
                        "default CPU thread count is 4 (include/qlten/framework/consts.h:24)",
    }, result


def cpu_model() -> str:
    """Model name of the host CPU (first entry of /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def run_e2e(args) -> int:
    """End-to-end two-site update on the U(1)xU(1) D=4096 complex128 tensors
    made Hermitian: 20 Lanczos steps with the matvec in split-BF16 mode and
    FP64 reorthogonalisation, then the truncated block SVD of the ground
    state over (l, s1 | s2, r) at cutoff 1e-10, Dmax = 4096.  The same
    computation in full FP64 on the CPU (oracle matvec, numpy LAPACK SVD) is
    the reference for the eigenvalue and truncation-weight errors."""
    from tensortoolkit_amd import decompose as DEC, lanczos as LZ
    capi.init(0)
    stream = C.c_void_p()
    capi.check(capi.lib().qlt_stream_create(C.byref(stream)))
    t0 = time.perf_counter()
    mv = W.config4()
    print(f"[e2e] tensors built in {time.perf_counter() - t0:.1f}s, psi {mv.psi.raw_size} elements",
          file=sys.stderr, flush=True)
    steps, cutoff, dmax = args.lanczos_steps, 1e-10, 4096
    mode = {"split": capi.QLT_MODE_SPLIT_BF16, "split_i8": capi.QLT_MODE_SPLIT_I8,
            "native": capi.QLT_MODE_NATIVE}[args.e2e_mode]
    hv = LZ.DeviceMatvec(mv, mode, stream)
    warm_in = capi.DeviceBuffer.from_numpy(mv.psi.data)
    warm_out = capi.DeviceBuffer(mv.psi.raw_size * 16)
    hv.apply(warm_in.ptr, warm_out.ptr)          # warm-up (first-launch costs)
    capi.check(capi.lib().qlt_stream_sync(stream))
    lz_loop = LZ.lanczos(hv, mv.psi.data.copy(), steps)
    print(f"[e2e] gpu lanczos, host loop: {lz_loop.seconds:.3f}s ({lz_loop.matvec_seconds:.3f}s in matvecs), "
          f"eigenvalue {lz_loop.eigenvalue!r}", file=sys.stderr, flush=True)
    arena = LZ.KrylovArena(hv, steps)
    LZ.lanczos_graph(hv, mv.psi.data.copy(), steps, arena=arena)    # warm-up of the captured iteration
    lz = LZ.lanczos_graph(hv, mv.psi.data.copy(), steps, arena=arena)
    print(f"[e2e] gpu lanczos, one hipGraph per iteration replayed: {lz.seconds:.3f}s "
          f"({lz.matvec_seconds:.3f}s replaying), eigenvalue {lz.eigenvalue!r}", file=sys.stderr, flush=True)
    t1 = time.perf_counter()
    res = DEC.svd(mv.psi, 2, (0, 0), cutoff, 1, dmax, t_dev=lz.ground_state, stream=stream, to_host=False)
    svd_s = time.perf_counter() - t1
    print(f"[e2e] gpu svd: {svd_s:.3f}s, D {res.D}, trunc err {res.trunc_err!r}, sweeps {res.sweeps}",
          file=sys.stderr, flush=True)
    line = {"metric": "two-site update: 20-step Lanczos (split-BF16 matvec, FP64 orthogonalisation) + FP64 block SVD",
            "workload": mv.name, "lanczos_steps": steps, "matvec_mode": args.e2e_mode,
            "gpu": {"eigenvalue": lz.eigenvalue, "trunc_err": res.trunc_err, "D": res.D,
                    "lanczos_s": round(lz.seconds, 4), "lanczos_replay_s": round(lz.matvec_seconds, 4),
                    "lanczos_host_loop_s": round(lz_loop.seconds, 4),
                    "svd_s": round(svd_s, 4), "svd_sweeps": res.sweeps,
                    "svd_block_matrices": len(res.svals)}}
    if not args.no_cpu_baseline:
        from oracle import decomp as OD, lanczos as OL
        from oracle import oracle as O_
        cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores) or cores))
        O_.use_openblas(cores)
        theta, gs, _, _, cpu_lz_s = OL.lanczos(mv, mv.psi.data.copy(), steps)
        psi = mv.psi
        psi_cpu = type(psi)(psi.indexes, psi.coords, psi.dims, dtype=psi.dtype, data=gs, offsets=psi.offsets)
        t2 = time.perf_counter()
        _, err_cpu, D_cpu, _ = OD.svd(psi_cpu, 2, cutoff, 1, dmax)
        cpu_svd_s = time.perf_counter() - t2
        line["cpu_fp64"] = {"eigenvalue": theta, "trunc_err": err_cpu, "D": D_cpu,
                            "lanczos_s": round(cpu_lz_s, 3), "svd_s": round(cpu_svd_s, 3), "cores": cores}
        line["eigenvalue_rel_err"] = abs(lz.eigenvalue - theta) / abs(theta)
        line["trunc_err_abs_diff"] = abs(res.trunc_err - err_cpu)
        line["D_equal"] = res.D == D_cpu
    print(json.dumps(line), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--settle-ms", type=float, default=200.0,
                    help="untimed run of each mode before its warmup steps, to bring the GPU clocks up")
    ap.add_argument("--config", default="u1u1_D4096", choices=["u1u1_D4096", "u1_D2048", "u1_lognormal"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--mode", default="all", choices=["all", "both", "native", "split", "split_bf16", "split_i8"],
                    help="contraction modes to time: all = native, split_bf16 and split_i8; both = native and "
                         "split_bf16; split = both split modes (the JSON line needs native for split_vs_native)")
    ap.add_argument("--graph-replays", type=int, default=100)
    ap.add_argument("--autotune", action="store_true",
                    help="run the autotuner on the chain per mode first (untimed) and time under its result; "
                         "default: the built-in tile rules")
    ap.add_argument("--no-config2", action="store_true",
                    help="skip the lognormal-degeneracy FP64 chain (autotuned, hipGraph replayed) reported beside "
                         "the default workload")
    ap.add_argument("--calibration-copy-mb", type=int, default=0,
                    help="run one device-to-device copy of this size first (scales rocprofv3 "
                         "FETCH_SIZE / WRITE_SIZE counters to bytes)")
    ap.add_argument("--e2e", action="store_true",
                    help="run the end-to-end two-site update (Lanczos + block SVD) instead of the Contract benchmark")
    ap.add_argument("--e2e-mode", default="split", choices=["split", "split_i8", "native"])
    ap.add_argument("--lanczos-steps", type=int, default=20)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported in this build "
              "(the contraction would shard by output block across GPUs)", file=sys.stderr)
        return 2

    if args.e2e:
        return run_e2e(args)
    capi.init(0)
    if args.calibration_copy_mb:
        nbytes = args.calibration_copy_mb << 20
        src, dst = capi.DeviceBuffer(nbytes), capi.DeviceBuffer(nbytes)
        capi.check(capi.lib().qlt_memset_zero(src.ptr, nbytes, None))
        capi.check(capi.lib().qlt_memcpy_d2d(dst.ptr, src.ptr, nbytes))
        capi.check(capi.lib().qlt_device_sync())
        print(f"[bench] calibration copy: {nbytes} bytes device to device", file=sys.stderr, flush=True)
        src.close()
        dst.close()
    stream = C.c_void_p()
    capi.check(capi.lib().qlt_stream_create(C.byref(stream)))
    hip = Hip()
    t0 = time.perf_counter()
    mv = {"u1u1_D4096": W.config3, "u1_D2048": W.config1, "u1_lognormal": W.config2}[args.config]()
    gen_s = time.perf_counter() - t0
    runner = MatvecRunner(mv, stream)
    flops_step = float(sum(runner.flops))
    print(f"[bench] {mv.name}: tensors built in {gen_s:.1f}s, plans in {runner.plan_seconds:.2f}s, "
          f"tasks {[int(p.info.n_tasks) for p in runner.plans]}, GFLOP/step {flops_step / 1e9:.2f}",
          file=sys.stderr, flush=True)

    results = {}
    modes = [("native", capi.QLT_MODE_NATIVE), ("split_bf16", capi.QLT_MODE_SPLIT_BF16),
             ("split_i8", capi.QLT_MODE_SPLIT_I8)]
    if args.mode == "both":
        modes = modes[:2]
    elif args.mode != "all":
        modes = [m for m in modes if m[0].startswith(args.mode)]
    outputs = {}
    dom = int(np.argmax(runner.flops))
    for name, mode in modes:
        tune = None
        TUNE.set_tuning(None)
        if args.autotune:
            t_tune = time.perf_counter()
            res = TUNE.autotune(runner.tune_items(), runner.dtype, mode, reps=3, max_buckets=16, stream=stream)
            tune = {"trials": res.trials, "ms_before": round(res.ms_before, 4), "ms_after": round(res.ms_after, 4),
                    "overrides": TUNE.overrides(res.tuning), "seconds": round(time.perf_counter() - t_tune, 1)}
            print(f"[bench] {name}: autotune {res.trials} trials in {tune['seconds']} s, chain "
                  f"{res.ms_before:.3f} -> {res.ms_after:.3f} ms, {tune['overrides']} overrides",
                  file=sys.stderr, flush=True)
        runner.rebuild_plans()
        try:
            total_ms, wall, per_launch, kernel = time_mode(runner, hip, mode, args.steps, args.warmup, dom,
                                                           args.settle_ms * 1e-3)
        except capi.QltError as e:
            print(f"[bench] mode {name} unavailable: {e}", file=sys.stderr, flush=True)
            continue
        outputs[name] = runner.result()
        graph_ms = time_graph(runner, hip, mode, args.graph_replays) if args.graph_replays else None
        results[name] = dict(total_ms=total_ms, wall=wall, per_launch_ms=per_launch, graph_ms=graph_ms,
                             kernel=kernel, autotune=tune)
        print(f"[bench] {name}: {total_ms / args.steps:.3f} ms/step, per launch {np.round(per_launch, 3).tolist()}, "
              f"dominant kernel {kernel['name']} {kernel['ms']:.4f} ms"
              + (f", graph replay {graph_ms:.3f} ms/step" if graph_ms else ""), file=sys.stderr, flush=True)

    config2 = None
    if args.config == "u1u1_D4096" and not args.no_config2 and args.graph_replays:
        config2 = config2_graph_replay(hip, stream, args.graph_replays, tune=True)
        print(f"[bench] {config2['workload']}: graph replay {config2['ms_per_replay']:.3f} ms "
              f"({config2['gflops_replay']:.0f} GFLOP/s), eager {config2['ms_per_step_eager']:.3f} ms",
              file=sys.stderr, flush=True)

    # correctness of the step output against the FP64 CPU oracle
    errors = {}
    cpu = None
    ref = None
    if not args.no_cpu_baseline:
        cpu, ref = cpu_baseline(mv)
        out_t = runner.out_tensors[-1]
        for name, got in outputs.items():
            worst = 0.0
            for i in range(out_t.n_blks):
                sz = int(np.prod(out_t.dims[i]))
                o = out_t.offsets[i]
                den = np.linalg.norm(ref[o:o + sz])
                if den > 0:
                    worst = max(worst, float(np.linalg.norm(got[o:o + sz] - ref[o:o + sz]) / den))
            errors[name] = worst

    primary = next((m for m in ("split_bf16", "split_i8", "native") if m in results), None)
    if primary is None:
        print("[bench] no mode ran", file=sys.stderr, flush=True)
        return 1
    r = results[primary]
    ms_step = r["total_ms"] / args.steps
    value = flops_step / (ms_step * 1e-3) / 1e9
    roofline = kernel_roofline(r["kernel"], mv.name)

    def contraction_roofline(res):
        """The heaviest contraction as a whole: its tile-class kernels run
        concurrently on the class streams, so the unit that reaches the MFMA
        bound together is the contraction (flops over its event-timed launch)."""
        ms = float(res["per_launch_ms"][dom])
        tf = runner.flops[dom] / (ms * 1e-3) / 1e12
        return {"unit_of_work": f"contraction {dom} (all tile classes)", "ms_per_launch": round(ms, 5),
                "algorithmic_flops": runner.flops[dom], "bound": "mfma", "achieved": round(tf, 3),
                "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": round(tf / PEAK_FP64_TFLOPS, 4),
                "note": "FP64-equivalent rate" if res is not results.get("native") else "FP64"}
    native_gf = None
    if "native" in results:
        native_gf = flops_step / (results["native"]["total_ms"] / args.steps * 1e-3) / 1e9
    def split_i8_summary():
        """The int8-digit split mode beside the metric's split-BF16 one."""
        if "split_i8" not in results:
            return None
        res = results["split_i8"]
        gf = flops_step / (res["total_ms"] / args.steps * 1e-3) / 1e9
        return {"gflops": round(gf, 2), "ms_per_step": round(res["total_ms"] / args.steps, 4),
                "vs_native": round(gf / native_gf, 3) if native_gf else None,
                "roofline": kernel_roofline(res["kernel"], mv.name),
                "graph_replay_ms_per_step": round(res["graph_ms"], 4) if res["graph_ms"] else None}

    line = {
        "metric": "block-sparse Contract GFLOP/s, split-BF16 grouped vs FP64, at <=1e-12 rel err",
        "value": round(value, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "settle_ms": args.settle_ms,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128" if mv.dtype == "c128" else "f64",
        "data": "synthetic",
        "config": {"workload": mv.name, "path": primary, "contractions_per_step": 4,
                   "gflop_per_step": round(flops_step / 1e9, 3),
                   "tasks": [int(p.info.n_tasks) for p in runner.plans]},
        "roofline": roofline,
        "native_gflops": round(native_gf, 2) if native_gf else None,
        "native_roofline": kernel_roofline(results["native"]["kernel"], mv.name) if "native" in results else None,
        "contraction_roofline": {name: contraction_roofline(res) for name, res in results.items()},
        "gflops_plan_plus_execute": round(flops_step / (runner.plan_seconds + ms_step * 1e-3) / 1e9, 2),
        "split_vs_native": round(value / native_gf, 3) if (native_gf and primary != "native") else None,
        "max_rel_err_vs_oracle": {k: float(f"{v:.3e}") for k, v in errors.items()},
        "graph_replay_ms_per_step": {k: round(v["graph_ms"], 4) for k, v in results.items() if v["graph_ms"]},
        "plan_build_s": round(runner.plan_seconds, 3),
        "autotune": {k: v["autotune"] for k, v in results.items()},
        "split_i8": split_i8_summary(),
        "config2_graph_replay": config2,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

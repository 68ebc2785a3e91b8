"""Benchmark: 30-qubit depth-40 seeded random circuit on one MI355X.

One step = one full run of the depth-40 circuit (30 single-qubit gates from
{H, X, RX, RY, RZ} + 10 CNOT/CZ on disjoint pairs per layer, SmallRng seed
2024) applied in place to a resident complex128 state (16 GiB) through the
fused tile-pass plan.  The state is set to |0..0> once before the warmup.

Prints one JSON line: gate-layers/s, the roofline of the dominant kernel
(tile pass; algorithmic bytes = 2 * 2^n * 16 per launch, its average launch
duration from HIP events on the engine's stream) and the CPU baseline (the
oracle's restatement of the reference CPU path, timed on a bounded sample of
the same circuit on this host).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def inplace_stream_rate(torch, nbytes: int) -> float:
    """GB/s of a plain in-place elementwise multiply (torch's kernel) over
    `nbytes` of device memory: the same one-read-one-write pattern as a tile
    pass, i.e. the streaming rate this GPU reaches for it."""
    x = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda:0")
    x.mul_(1.0)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(5):
        x.mul_(1.0)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / 5
    del x
    torch.cuda.empty_cache()
    return 2 * nbytes / (ms * 1e-3) / 1e9


def measured_traffic(n, args, tile_bits, low, reg_bits, passes, compiled):
    """HBM bytes per tile-pass launch measured with rocprofv3 (FETCH_SIZE and
    WRITE_SIZE, copy-calibrated) for this exact workload, from
    profiles/hbm_traffic.json; None when the workload was not profiled."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        table = json.load(f)
    want = {"qubits": n, "depth": args.depth, "seed": args.seed, "precision": args.precision,
            "pass_kernels": "compiled" if compiled else ("register" if reg_bits else "lds"),
            "tile_qubits": tile_bits, "low_qubits": low, "register_bits": reg_bits, "passes": passes}
    for e in table.get("entries", []):
        if all(e.get(k) == v for k, v in want.items()):
            return e["bytes_per_launch"]
    return None


def cpu_baseline(circuit, qubits: int, sample_ops: int):
    """Times the restated reference CPU path (clone+collect 1q apply with
    OpenMP, sequential CNOT/CZ loops; oracle/qf_oracle.cpp) on the first
    `sample_ops` ops of layer 0 in the 3:1 single:two-qubit mix of a layer,
    and scales to layers/s."""
    import numpy as np

    from oracle import oracle as O

    ops = circuit.to_ops()
    one = [o for o in ops[: qubits + 10] if o["kind"] == 0]
    two = [o for o in ops[: qubits + 10] if o["kind"] != 0]
    n1 = max(1, (3 * sample_ops) // 4)
    n2 = max(1, sample_ops - n1)
    sample = np.array(one[:n1] + two[:n2], dtype=ops.dtype)
    t = O.time_run(qubits, sample)
    per_1q = None
    # Per-layer time: 30 one-qubit + 10 two-qubit ops, scaled from the sample mix.
    layer_time = t * (qubits + 10) / len(sample)
    return {
        "value": 1.0 / layer_time,
        "unit": "layers/s",
        "cores": O.max_threads(),
        "kind": "port",
        "sample": f"{n1} single-qubit + {n2} two-qubit ops of layer 0 at {qubits} qubits ({t:.1f} s), "
                  f"scaled to a {qubits}+10-op layer",
        "optimized_inplace": optimized_cpu(ops, qubits),
    }


def optimized_cpu(ops, qubits: int, sample_1q: int = 6, sample_2q: int = 2) -> dict:
    """An optimized CPU form beside the restatement (never in its place): the
    same per-pair arithmetic applied in place with OpenMP over the state
    (oracle or_fast_run), timed on 6 single-qubit and 2 two-qubit ops of layer
    0 after one untimed op that touches the state's pages, scaled to a layer."""
    import numpy as np

    from oracle import oracle as O

    one = [o for o in ops[: qubits + 10] if o["kind"] == 0]
    two = [o for o in ops[: qubits + 10] if o["kind"] != 0]
    amps = np.zeros(1 << qubits, dtype=np.complex128)
    amps[0] = 1.0
    O.fast_run(qubits, np.array(one[:1], dtype=ops.dtype), amps)
    sample = np.array(one[1:1 + sample_1q] + two[:sample_2q], dtype=ops.dtype)
    t0 = time.perf_counter()
    O.fast_run(qubits, sample, amps)
    t = time.perf_counter() - t0
    del amps
    layer = t / len(sample) * (qubits + 10)
    return {"value": 1.0 / layer, "unit": "layers/s", "cores": O.max_threads(),
            "sample": f"{sample_1q} single-qubit + {sample_2q} two-qubit ops of layer 0 at {qubits} qubits, in "
                      f"place ({t:.1f} s), scaled to a {qubits}+10-op layer"}


def vqe_bench(be, sets: int, shift_sets: int, with_cpu: bool) -> dict:
    """16-qubit hardware-efficient ansatz (8 layers of Ry, Rz on every qubit,
    linear CNOT ladder, 256 parameters ~ U[-pi, pi]) with a Heisenberg + Z-field
    chain: batched adjoint gradients for `sets` parameter sets against the
    plain parameter-shift GPU form (2 circuits per parameter) on a subset, in
    complex128 and complex64.  The roofline is the whole adjoint call's
    algorithmic bytes (forward passes 2 x 2^n x s per set, costate 2 x 2^n x s,
    reverse two-state passes 4 x 2^n x s) over its device time, with each
    phase's share from HIP events (qf_backend_last_gradient_profile)."""
    import numpy as np

    from quantumforge_amd import gradients as G
    from quantumforge_amd.workloads import hea_circuit, heisenberg_hamiltonian, uniform_parameters

    nq, layers = 16, 8
    pc = hea_circuit(nq, layers)
    H = heisenberg_hamiltonian(nq, seed=16)
    params = uniform_parameters(pc.parameter_count, sets, seed=4096)
    P = pc.parameter_count
    res = {"workload": f"{nq}q HEA {layers} layers ({P} params), Heisenberg+Z chain, {sets} sets"}
    ref_e = ref_g = None
    for precision in ("complex128", "complex64"):
        G.batch_adjoint_gradient(be, pc, nq, params[: min(sets, 64)], H, precision)  # compiles the plans
        be.set_profiling(False)
        # The first full-size call also grows the backend's device arena and
        # pinned staging; the timed call is the steady state of a training loop.
        t0 = time.perf_counter()
        G.batch_adjoint_gradient(be, pc, nq, params, H, precision)
        t_first = time.perf_counter() - t0
        t0 = time.perf_counter()
        e, g = G.batch_adjoint_gradient(be, pc, nq, params, H, precision)
        t_adj = time.perf_counter() - t0
        prof_plain = be.last_gradient_profile()
        be.set_profiling(True)
        G.batch_adjoint_gradient(be, pc, nq, params, H, precision)
        prof = be.last_gradient_profile()
        be.set_profiling(False)
        log(f"vqe {precision}: adjoint {sets} sets x {P} params in {t_adj * 1e3:.1f} ms (first call "
            f"{t_first * 1e3:.1f} ms); phases {prof}")
        k = max(1, min(shift_sets, sets))
        G.batch_shift_gradient(be, pc, nq, params[:1], H, precision)  # warmup
        t0 = time.perf_counter()
        es, gs = G.batch_shift_gradient(be, pc, nq, params[:k], H, precision)
        t_shift = time.perf_counter() - t0
        log(f"vqe {precision}: parameter shift {k} sets in {t_shift * 1e3:.1f} ms")
        adj_gps = sets * P / t_adj
        shift_gps = k * P / t_shift
        dev_ms = prof["forward_ms"] + prof["costate_ms"] + prof["reverse_ms"] + prof["finish_ms"]
        alg = prof["forward_bytes"] + prof["costate_bytes"] + prof["reverse_bytes"]

        def gbs(b, ms):
            return round(b / (ms * 1e-3) / 1e9, 1) if ms > 0 else None

        r = {
            "adjoint_grads_per_s": round(adj_gps, 1),
            "adjoint_ms": round(t_adj * 1e3, 2),
            "first_call_ms": round(t_first * 1e3, 2),
            "shift_grads_per_s": round(shift_gps, 1),
            "shift_sets_timed": k,
            "ratio_vs_shift": round(adj_gps / shift_gps, 2),
            "max_abs_adjoint_minus_shift": float(np.max(np.abs(g[:k] - gs))),
            "max_abs_energy_diff": float(np.max(np.abs(e[:k] - es))),
            "plan": {"forward_passes": prof["forward_passes"], "reverse_passes": prof["reverse_passes"],
                     "forward_kernels": prof_plain["forward_kernels"],
                     "reverse_kernels": prof_plain["reverse_kernels"], "costate_kernel": prof["costate_kernel"],
                     "waves": prof["waves"]},
            "roofline": {"bound": "hbm", "achieved": gbs(alg, dev_ms), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(alg / (dev_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if dev_ms > 0 else None,
                         "alg_bytes": alg, "device_ms": round(dev_ms, 3),
                         "phases": {
                             "forward": {"ms": round(prof["forward_ms"], 3),
                                         "GBs": gbs(prof["forward_bytes"], prof["forward_ms"])},
                             "costate": {"ms": round(prof["costate_ms"], 3),
                                         "GBs": gbs(prof["costate_bytes"], prof["costate_ms"])},
                             "reverse": {"ms": round(prof["reverse_ms"], 3),
                                         "GBs": gbs(prof["reverse_bytes"], prof["reverse_ms"])},
                             "finish_ms": round(prof["finish_ms"], 3)}},
        }
        if precision == "complex128":
            ref_e, ref_g = e, g
        else:
            # complex64 amplitudes, FP64 reductions: error against the complex128
            # run (itself within 1e-14 of the FP64 oracle, below).
            r["max_abs_energy_err_vs_c128"] = float(np.max(np.abs(e - ref_e)))
            r["max_abs_grad_err_vs_c128"] = float(np.max(np.abs(g - ref_g)))
        res[precision] = r
    if with_cpu:
        from oracle import oracle as O

        t0 = time.perf_counter()
        gc = O.pc_adjoint_gradient(pc.to_pops(), P, 0, params[0], [], nq, H.to_terms())
        t_cpu = time.perf_counter() - t0
        ec = O.pc_expectation(pc.to_pops(), P, 0, params[0], [], nq, H.to_terms())
        res["cpu_adjoint_grads_per_s"] = round(P / t_cpu, 1)
        res["cpu_sample"] = "1 set, oracle restatement of adjoint_gradient_with_inputs (store-every-state)"
        res["max_abs_gpu_vs_cpu_adjoint"] = float(np.max(np.abs(ref_g[0] - gc)))
        res["abs_gpu_vs_cpu_energy"] = float(abs(ref_e[0] - ec))
    # the headline figures of the section (complex128)
    for key in ("adjoint_grads_per_s", "adjoint_ms", "shift_grads_per_s", "ratio_vs_shift", "roofline"):
        res[key] = res["complex128"][key]
    return res


def circuit_section(be, torch, circuit, n: int, precision: str, steps: int, warmup: int, label: str,
                    kernel: int = 0, max_block: int = 0, stream_probe: bool = False) -> dict:
    """Times `steps` runs of `circuit` on a resident |0..0> state with its
    default plan (or the plan of `kernel`; HIP events on the engine's stream)."""
    from quantumforge_amd.backend import Plan

    plan = Plan(circuit, n, precision=precision, kernel=kernel, max_block=max_block)
    stats = plan.stats()
    compiled = bool(plan.describe().get("compiled", 0))
    if compiled:
        plan.compile(precision)
    st = be.zero_state(n, precision=precision)
    stream = torch.cuda.ExternalStream(be.stream())
    for _ in range(warmup):
        st.run_circuit(circuit, plan)
    be.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(steps):
        st.run_circuit(circuit, plan)
    b.record(stream)
    be.synchronize()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / steps
    amp = 16 if precision == "complex128" else 8
    per_pass = ms / max(1, stats["passes"])
    res = {"workload": label, "qubits": n, "precision": precision, "ops": stats["ops_in"],
           "passes": stats["passes"], "pass_kernels": "compiled" if compiled else "generic",
           "kernels_run": plan.kernels_run(), "dense_blocks": stats["dense_blocks"],
           "ms_per_circuit": round(ms, 3), "ms_per_pass": round(per_pass, 4),
           "GBs_per_pass": round(2 * (1 << n) * amp / (per_pass * 1e-3) / 1e9, 1),
           "norm_after": st.norm_sqr(), "steps": steps, "warmup": warmup}
    if stream_probe:
        # the device's one-read-one-write rate over this state's bytes
        res["stream_kernels_GBs"] = {name: round(st.stream_rate(v, 3)[0], 1)
                                     for v, name in ((1, "nontemporal"), (2, "four_words"), (4, "eight_words"))}
    del st
    return res


MFMA_PEAK_TFLOPS = {"complex128": 78.6, "complex64": 157.3}  # dense FP64 / FP32 matrix peaks


def qft_like_bench(be, torch, steps: int) -> list:
    """33 qubits complex128 and 34 qubits complex64 (128 GiB each): 2 layers of
    H on every qubit, exact controlled phases CP(pi/2^d) at distances d <= 8,
    RZ(2 pi u) on every qubit (SmallRng seed 33).  Each circuit runs through
    the compiled tile passes (the default) and, for comparison, through dense
    5-qubit blocks on the matrix cores (plan kernel = 4), whose MFMA
    utilisation is the block products' flops (8 x 2^5 per amplitude and pass)
    over the pass time and the dense matrix peak of the precision."""
    from quantumforge_amd.workloads import qft_like_circuit

    out = []
    for n, prec in ((33, "complex128"), (34, "complex64")):
        _, native = qft_like_circuit(n, 2, seed=33)
        r = circuit_section(be, torch, native, n, prec, steps, 1,
                            f"QFT-like, {n} qubits, 2 layers, exact controlled phases, {prec}", stream_probe=True)
        r["layers_per_s"] = round(2 / (r["ms_per_circuit"] * 1e-3), 3)
        log(f"qft-like {n}q {prec}: {r['ms_per_circuit']:.1f} ms/circuit, {r['passes']} passes, "
            f"{r['GBs_per_pass']:.0f} GB/s per pass")
        d = circuit_section(be, torch, native, n, prec, steps, 1,
                            f"QFT-like, {n} qubits, 2 layers, dense 5-qubit blocks on the matrix cores, {prec}",
                            kernel=4, max_block=5)
        flops = d["passes"] * 8 * 32 * float(1 << n)
        tflops = flops / (d["ms_per_circuit"] * 1e-3) / 1e12
        d["layers_per_s"] = round(2 / (d["ms_per_circuit"] * 1e-3), 3)
        d["mfma"] = {"block_qubits": 5, "tflops": round(tflops, 2), "peak_tflops": MFMA_PEAK_TFLOPS[prec],
                     "utilisation": round(tflops / MFMA_PEAK_TFLOPS[prec], 4)}
        log(f"qft-like {n}q {prec} dense blocks: {d['ms_per_circuit']:.1f} ms/circuit, {d['passes']} passes, "
            f"{d['GBs_per_pass']:.0f} GB/s per pass, MFMA {tflops:.1f} TFLOP/s")
        r["dense_blocks_ab"] = d
        r["faster"] = "compiled tile passes" if r["ms_per_circuit"] <= d["ms_per_circuit"] else "dense MFMA blocks"
        out.append(r)
    return out


def uneven_bench(be, count: int, seed: int, check_instances: int) -> dict:
    """QAOA MaxCut adjoint gradients of `count` instances of different widths
    (even n in 18..26, random 3-regular graphs with U[0.5, 1) weights, p = 6,
    gamma, beta ~ U[0, pi]) through the uneven-batch engine; energies and
    gradients of the `check_instances` narrowest instances against the FP64
    oracle (the reference's store-every-state adjoint)."""
    from quantumforge_amd import gradients as G
    from quantumforge_amd.workloads import (maxcut_cost_hamiltonian, qaoa_occurrence_angles, qaoa_parameterized,
                                            qaoa_uneven_instances)

    specs = qaoa_uneven_instances(count, seed=seed, p=6)
    insts = []
    for sp in specs:
        prob = sp["problem"]
        pc = qaoa_parameterized(prob.num_qubits, prob.edges, sp["p"])
        ang = qaoa_occurrence_angles(prob.edges, prob.num_qubits, sp["p"], sp["gammas"], sp["betas"])
        insts.append((pc, prob.num_qubits, ang, maxcut_cost_hamiltonian(prob)))
    # The first call grows the backend's device arena (a caching allocator,
    # about 230 GB here); the timed call is the steady state of a loop over
    # such batches.  The first call's time is reported beside it: in a
    # process started just after another one freed a large part of HBM, the
    # allocation alone has taken seconds.
    t0 = time.perf_counter()
    G.uneven_adjoint_gradient(be, insts)
    t_first = time.perf_counter() - t0
    t0 = time.perf_counter()
    e, g, stats = G.uneven_adjoint_gradient(be, insts)
    dt = time.perf_counter() - t0
    # The rest of the bench gets the memory back.
    be.release_cached()
    occ = sum(pc.parameter_count for pc, _, _, _ in insts)
    log(f"uneven: {count} instances, {occ} rotation occurrences, {dt:.2f} s (first call {t_first:.2f} s), "
        f"stats {stats}")
    dev_s = stats["device_ms"] * 1e-3
    res = {"workload": f"{count} QAOA MaxCut instances, even n in 18..26, p=6, complex128, seed {seed}",
           "seconds": round(dt, 3), "first_call_s": round(t_first, 3), "instances_per_s": round(count / dt, 2),
           "occurrence_grads_per_s": round(occ / dt, 1), "stats": stats,
           "mean_qubits": float(np.mean([q for _, q, _, _ in insts])),
           "roofline": {"bound": "hbm", "achieved": round(stats["alg_bytes"] / dev_s / 1e9, 1) if dev_s else None,
                        "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(stats["alg_bytes"] / dev_s / 1e9 / HBM_PEAK_GBS, 4) if dev_s else None,
                        "alg_bytes": stats["alg_bytes"], "device_s": round(dev_s, 3)}}
    if check_instances > 0:
        from oracle import oracle as O

        order = sorted(range(count), key=lambda k: insts[k][1])[:check_instances]
        errs = []
        t0 = time.perf_counter()
        for k in order:
            pc, q, ang, H = insts[k]
            t = H.to_terms()
            we = O.pc_expectation(pc.to_pops(), pc.parameter_count, 0, ang, [], q, t)
            wg = O.pc_adjoint_gradient(pc.to_pops(), pc.parameter_count, 0, ang, [], q, t)
            errs.append({"instance": k, "qubits": q, "energy_err": float(abs(e[k] - we)),
                         "max_grad_err": float(np.max(np.abs(g[k] - wg)))})
        res["fp64_check"] = {"instances": errs, "cpu_s": round(time.perf_counter() - t0, 1),
                             "max_energy_err": max(x["energy_err"] for x in errs),
                             "max_grad_err": max(x["max_grad_err"] for x in errs)}
        log(f"uneven: FP64 check of {len(errs)} instances: {res['fp64_check']}")
    return res


def complex64_section(be, torch, circuit, n: int, steps: int, warmup: int) -> dict:
    """The headline circuit in complex64 (8 GiB at 30 qubits): layers/s and the
    pass roofline (2 x 2^n x 8 bytes per launch), and the fidelity of the
    complex64 state against the complex128 state of the same circuit, both
    computed on the device (the complex64 state widened exactly to
    complex128, FP64 compensated inner product)."""
    r = circuit_section(be, torch, circuit, n, "complex64", steps, warmup,
                        f"random layered circuit, {n} qubits, depth 40, complex64, fused tile passes")
    r["layers_per_s"] = round(40 / (r["ms_per_circuit"] * 1e-3), 3)
    r["roofline"] = {"bound": "hbm", "achieved": r["GBs_per_pass"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(r["GBs_per_pass"] / HBM_PEAK_GBS, 4), "bytes_per_launch": 2 * (1 << n) * 8}
    from quantumforge_amd.backend import Plan

    a = be.zero_state(n)
    a.run_circuit(circuit, Plan(circuit, n))
    b = be.zero_state(n, precision="complex64")
    b.run_circuit(circuit, Plan(circuit, n, precision="complex64"))
    wide = be.zero_state(n)
    wide.copy_from_device(b)
    del b
    ov = a.inner(wide)
    nb = wide.norm_sqr()
    del a, wide
    r["fidelity_vs_complex128"] = abs(ov) ** 2 / nb
    r["one_minus_fidelity"] = 1.0 - r["fidelity_vs_complex128"]
    r["norm_sqr_complex64"] = nb
    log(f"complex64 {n}q: {r['ms_per_circuit']:.1f} ms/circuit, fidelity vs complex128 1 - {r['one_minus_fidelity']:.3e}")
    return r


def fidelity_oracle_section(be, n: int, seed: int) -> dict:
    """The headline circuit family at `n` qubits (depth 40) against the FP64
    CPU oracle (its in-place OpenMP run, the reference's per-pair arithmetic):
    complex128 and complex64 states from the compiled passes, max amplitude
    error and fidelity |<oracle|gpu>|^2 over the whole state."""
    from oracle import oracle as O
    from quantumforge_amd.backend import Plan
    from quantumforge_amd.workloads import random_layer_circuit

    c = random_layer_circuit(n, 40, seed=seed)
    st = be.zero_state(n)
    plan = Plan(c, n)
    st.run_circuit(c, plan)
    got128 = st.download().amplitudes
    del st
    st = be.zero_state(n, precision="complex64")
    st.run_circuit(c, Plan(c, n, precision="complex64"))
    got64 = st.download_c64().astype(np.complex128)
    del st
    t0 = time.perf_counter()
    want = O.fast_run(n, c.to_ops())
    t = time.perf_counter() - t0
    res = {"workload": f"random layered circuit, {n} qubits, depth 40, seed {seed}, vs the FP64 CPU oracle",
           "oracle_s": round(t, 1), "oracle_threads": O.max_threads()}
    for name, got in (("complex128", got128), ("complex64", got64)):
        err = float(np.max(np.abs(got - want)))
        fid = float(abs(np.vdot(want, got)) ** 2 / np.vdot(got, got).real)
        res[name] = {"max_abs_amp_err": err, "fidelity": fid, "one_minus_fidelity": 1.0 - fid}
    log(f"fidelity {n}q vs oracle: {res}")
    return res


def pcie_bench(be, qubits: int) -> dict:
    """Host<->device rate of the boundary's host-buffer entry points
    (qf_state_upload / qf_state_download, pageable NumPy memory) for one
    complex128 state of `qubits` qubits."""
    from quantumforge_amd import StateVector

    amps = np.zeros(1 << qubits, dtype=np.complex128)
    amps[0] = 1.0
    sv = StateVector(amps, qubits)
    dev = be.upload(sv)  # warm the path
    be.synchronize()
    t0 = time.perf_counter()
    dev = be.upload(sv)
    be.synchronize()
    t_up = time.perf_counter() - t0
    t0 = time.perf_counter()
    out = dev.download()
    t_down = time.perf_counter() - t0
    assert out.amplitudes[0] == 1.0
    nbytes = amps.nbytes
    return {"qubits": qubits, "bytes": nbytes, "upload_GBs": round(nbytes / t_up / 1e9, 2),
            "download_GBs": round(nbytes / t_down / 1e9, 2), "memory": "pageable host (NumPy)"}


def density_bench(be, torch, n: int, layers: int, steps: int, with_cpu: bool) -> dict:
    """A density matrix of n qubits (row-major entries = a 2n-qubit state,
    4^n * 16 bytes resident) under `layers` noisy layers
    (workloads.noisy_layer_circuit, seed 7): compiled tile passes over the
    2n-qubit view, gates and channel superoperators fused.  One step = the
    whole circuit on the resident matrix, timed with HIP events on the
    engine's stream."""
    from quantumforge_amd.backend import Plan
    from quantumforge_amd.workloads import noisy_layer_circuit

    c = noisy_layer_circuit(n, layers, seed=7)
    plan = Plan.density(c, n)
    stats = plan.stats()
    desc = plan.describe()
    compiled = bool(desc.get("compiled", 0))
    info = plan.compile() if compiled else None
    dev = be.density_from_state(be.zero_state(n))
    stream = torch.cuda.ExternalStream(be.stream())
    dev.run_circuit(c, plan)
    be.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(steps):
        dev.run_circuit(c, plan)
    b.record(stream)
    be.synchronize()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / steps
    trace, purity = dev.trace(), dev.purity()
    del dev
    nbytes = (1 << (2 * n)) * 16
    per_pass = ms / max(1, stats["passes"])
    achieved = 2 * nbytes / (per_pass * 1e-3) / 1e9
    res = {"workload": f"density matrix, {n} qubits ({2 * n}-qubit view, {nbytes >> 30} GiB), {layers} noisy layers "
                       f"(rotations, CNOT/CZ, one channel per qubit and layer), complex128",
           "qubits": n, "layers": layers, "ops": len(c), "passes": stats["passes"],
           "pass_kernels": "compiled" if compiled else "generic",
           "compile_ms": round(info["compile_ms"], 1) if info else None,
           "ms_per_circuit": round(ms, 3), "layers_per_s": round(layers / (ms * 1e-3), 3),
           "ms_per_pass": round(per_pass, 4),
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "bytes_per_launch": 2 * nbytes},
           "trace_after": [trace.real, trace.imag], "purity_after": purity, "steps": steps}
    p0 = desc["passes"][0]
    want = {"workload": "density", "density_qubits": n, "layers": layers, "seed": 7, "precision": "complex128",
            "pass_kernels": "compiled" if compiled else "generic", "tile_qubits": p0["tile_bits"],
            "low_qubits": p0.get("low", 0), "register_bits": p0.get("reg_bits", 0), "passes": stats["passes"]}
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    res["roofline"]["traffic"] = None
    if os.path.exists(path):
        with open(path) as f:
            for e in json.load(f).get("entries", []):
                if all(e.get(k) == v for k, v in want.items()):
                    res["roofline"]["traffic"] = e["bytes_per_launch"]
    log(f"density {n}q: {ms:.1f} ms per {layers}-layer circuit, {stats['passes']} passes, "
        f"{achieved:.0f} GB/s per pass")
    if with_cpu:
        from oracle import oracle as O

        m = min(n, 11)
        c1 = noisy_layer_circuit(m, 1, seed=7)
        psi = np.zeros(1 << m, dtype=np.complex128)
        psi[0] = 1.0
        d = O.Density(m, state=psi)
        t0 = time.perf_counter()
        d.run(c1.to_ops())
        t = time.perf_counter() - t0
        scale = float(4 ** (n - m)) * (len(c) / layers) / len(c1)
        res["cpu_baseline"] = {
            "value": round(1.0 / (t * scale), 6), "unit": "layers/s", "cores": 1, "kind": "port",
            "sample": f"one noisy layer of {len(c1)} ops on a {m}-qubit density matrix ({t:.1f} s; the reference's "
                      f"DensityMatrix loops are sequential), scaled by the entry count 4^{n - m} to {n} qubits"}
    return res


def trajectory_bench(be, n: int, shots: int, layers: int, with_cpu: bool) -> dict:
    """Batched noisy shots (Kraus trajectories): `shots` trajectories of an
    n-qubit state through `layers` noisy layers, then every qubit measured,
    seed 11; one call of run_circuit_shots_batched, wall clock."""
    from quantumforge_amd import StateVector
    from quantumforge_amd.workloads import noisy_layer_circuit

    c = noisy_layer_circuit(n, layers, seed=11)
    for q in range(n):
        c.measure(q, q)
    init = StateVector.zero(n)
    be.run_circuit_shots_batched(init, c, 8, 1)  # warm the plans
    be.synchronize()
    t0 = time.perf_counter()
    r = be.run_circuit_shots_batched(init, c, shots, 11)
    dt = time.perf_counter() - t0
    res = {"workload": f"{shots} noisy trajectories, {n} qubits, {layers} noisy layers, measure all, complex128",
           "qubits": n, "shots": shots, "ops": len(c), "seconds": round(dt, 3), "shots_per_s": round(shots / dt, 1),
           "distinct_outcomes": len(r.counts)}
    log(f"trajectories {n}q: {shots} shots in {dt:.2f} s")
    if with_cpu:
        from oracle import oracle as O

        k = 64
        t0 = time.perf_counter()
        O.run_shots(init.amplitudes, n, c.to_ops(), k, 11, noise=True)
        t = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": round(k / t, 3), "unit": "shots/s", "cores": O.max_threads(),
                               "kind": "port", "sample": f"{k} shots of the same circuit ({t:.1f} s)"}
    return res


def lanczos_bench(be, sites: int, csr_sites: int, with_cpu: bool) -> dict:
    """The DMRG TFIM Lanczos on the device, matrix-free (`sites` sites,
    2^sites doubles per Krylov vector, 30 steps from 24 sites, 100 below),
    and sparse_lanczos_gpu on the same chain stored as CSR (`csr_sites`
    sites, 100 steps).  Wall clock of one call each.  Algorithmic bytes of step
    j: the operator's vectors (read v_j and v_{j-1}, write w: 3 x 8 B per
    entry), one reorthogonalization pass (read the j + 1 basis vectors twice,
    w three times) and the division (2 x 8 B)."""
    from quantumforge_amd import SparseOperator

    def model_bytes(dim, steps):
        return sum((3 + 2 * (j + 1) + 3 + 2) * 8 * dim for j in range(steps))

    be.tfim_lanczos(12)  # warm the kernels
    res = {}
    if sites > 0:
        t0 = time.perf_counter()
        e, a, b, m = be.tfim_lanczos(sites, details=True)
        dt = time.perf_counter() - t0
        dim = 1 << sites
        steps = int(np.count_nonzero(a[:m])) or m
        ach = model_bytes(dim, steps) / dt / 1e9
        res["tfim"] = {"workload": f"TFIM chain, {sites} sites (2^{sites} doubles per vector), {steps} Lanczos steps, "
                                   f"matrix-free", "sites": sites, "steps": steps, "seconds": round(dt, 4),
                       "ms_per_step": round(dt / steps * 1e3, 3), "lowest": e,
                       "model_GBs": round(ach, 1), "model_frac_of_8TBs": round(ach / HBM_PEAK_GBS, 4)}
        log(f"lanczos tfim {sites}: {dt:.3f} s for {steps} steps, E0 = {e:.12f}, {ach:.0f} GB/s (model bytes)")
    if csr_sites > 0:
        op = SparseOperator.tfim(csr_sites)
        t0 = time.perf_counter()
        e, a, b, size = be.sparse_lanczos(op, 100, 1e-10, details=True)
        dt = time.perf_counter() - t0
        res["csr"] = {"workload": f"TFIM chain as CSR, {csr_sites} sites, {op.nonzero_count()} stored values, "
                                  f"{size} Lanczos steps", "steps": size, "seconds": round(dt, 4),
                      "ms_per_step": round(dt / size * 1e3, 3), "lowest": e}
        log(f"lanczos csr {csr_sites}: {dt:.3f} s for {size} steps, E0 = {e:.12f}")
    if with_cpu and sites > 0:
        from oracle import oracle as O

        n = min(sites, 18)
        t0 = time.perf_counter()
        O.tfim_lanczos(n)
        t = time.perf_counter() - t0
        m_cpu = min(100, 1 << n)
        # Work of the restated loops: per step (n + 1) operator terms and 2 (j + 1) + 4 vector passes per entry.
        work = lambda nn, mm: sum((1 << nn) * ((nn + 1) + 2 * (j + 1) + 4) for j in range(mm))
        scale = work(sites, res["tfim"]["steps"]) / work(n, m_cpu)
        res["cpu_baseline"] = {"value": round(1.0 / (t * scale), 6), "unit": "TFIM Lanczos runs/s", "cores": 1,
                               "kind": "port",
                               "sample": f"the oracle's restatement (sequential, as DMRG::run_sparse_lanczos) at "
                                         f"{n} sites, {m_cpu} steps ({t:.1f} s), scaled by the loop work to "
                                         f"{sites} sites"}
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--depth", type=int, default=40)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--precision", default="complex128", choices=["complex128", "complex64"])
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--low", type=int, default=0)
    ap.add_argument("--kernel", type=int, default=0,
                    help="0 default, 1 LDS op-by-op tile kernel, 2 register kernel, 3 compiled passes")
    ap.add_argument("--reg-bits", type=int, default=0)
    ap.add_argument("--cpu-sample-ops", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--vqe-sets", type=int, default=4096, help="parameter sets of the 16-qubit VQE adjoint (0 = skip)")
    ap.add_argument("--vqe-shift-sets", type=int, default=4, help="sets timed with the parameter-shift GPU form")
    ap.add_argument("--qft-steps", type=int, default=2, help="timed runs of the 33q/34q QFT-like circuits (0 = skip)")
    ap.add_argument("--small-steps", type=int, default=20, help="timed runs of the 20-qubit circuit (0 = skip)")
    ap.add_argument("--pcie-qubits", type=int, default=26, help="state size of the upload/download timing (0 = skip)")
    ap.add_argument("--density-qubits", type=int, default=15, help="density-matrix section size (0 = skip)")
    ap.add_argument("--density-layers", type=int, default=4)
    ap.add_argument("--density-steps", type=int, default=2)
    ap.add_argument("--traj-qubits", type=int, default=16, help="noisy-trajectory section size (0 = skip)")
    ap.add_argument("--traj-shots", type=int, default=4096)
    ap.add_argument("--lanczos-sites", type=int, default=26, help="matrix-free TFIM Lanczos size (0 = skip)")
    ap.add_argument("--lanczos-csr-sites", type=int, default=22, help="CSR sparse Lanczos size (0 = skip)")
    ap.add_argument("--uneven-instances", type=int, default=512,
                    help="uneven QAOA adjoint batch size (0 = skip; the full workload is 512)")
    ap.add_argument("--uneven-check", type=int, default=2,
                    help="uneven instances (the narrowest) checked against the FP64 oracle")
    ap.add_argument("--c64-steps", type=int, default=3, help="timed runs of the complex64 headline circuit (0 = skip)")
    ap.add_argument("--fidelity-qubits", type=int, default=28,
                    help="width of the headline-family check against the FP64 CPU oracle (0 = skip)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (3 launches) to scale HBM counters")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the state-vector workload runs on a single MI355X; "
              "multi-GPU sharding is not part of this build (run with --gpus 1).", file=sys.stderr)
        return 2

    import torch

    if args.copy_calibration_mib > 0:
        nbytes = args.copy_calibration_mib << 20
        src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration": {"bytes_read": nbytes, "bytes_written": nbytes, "launches": 3}}))
        return 0

    from quantumforge_amd import StateVector
    from quantumforge_amd.backend import HipBackend, Plan
    from quantumforge_amd.workloads import random_layer_circuit

    n = args.qubits
    circuit = random_layer_circuit(n, args.depth, seed=args.seed)
    plan = Plan(circuit, n, fuse=True, tile_qubits=args.tile, low_qubits=args.low, kernel=args.kernel,
                reg_bits=args.reg_bits, precision=args.precision)
    stats = plan.stats()
    desc = plan.describe()
    tile_bits = desc["passes"][0]["tile_bits"]
    reg_bits = desc["passes"][0].get("reg_bits", 0)
    compiled = bool(desc.get("compiled", 0))
    log(f"plan: {stats['passes']} passes, {stats['phases']} register phases for {stats['ops_in']} ops "
        f"(tile {tile_bits} qubits, register bits {reg_bits}, {'compiled' if compiled else 'generic'} kernels)")
    compile_info = None
    if compiled:
        compile_info = plan.compile(args.precision)
        log(f"compiled {stats['passes']} pass kernels in {compile_info['compile_ms']:.0f} ms"
            f"{' (on-disk cache)' if compile_info['from_cache'] else ''}")

    be = HipBackend(0)
    state = be.zero_state(n, precision=args.precision)
    amp_bytes = 16 if args.precision == "complex128" else 8
    stream = torch.cuda.ExternalStream(be.stream())

    for w in range(args.warmup):
        state.run_circuit(circuit, plan)
        be.synchronize()
        log(f"warmup {w + 1}/{args.warmup} done")

    torch.cuda.synchronize()
    be.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    t0 = time.perf_counter()
    for s in range(args.steps):
        state.run_circuit(circuit, plan)
    end.record(stream)
    be.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = start.elapsed_time(end)
    log(f"timed {args.steps} steps: {gpu_ms:.1f} ms GPU, {wall * 1e3:.1f} ms wall")
    ran = plan.kernels_run()
    if args.kernel in (0, 3) and compiled and ran != ["compiled"]:
        raise RuntimeError(f"the headline plan was to run compiled passes but launched {ran}")

    norm = state.norm_sqr()
    log(f"norm after {args.warmup + args.steps} runs: {norm:.15f}")
    # The one-read-one-write ceiling of this device over the state's bytes:
    # hand-written streaming kernels (16-byte words; nontemporal; 64 and 128
    # bytes in flight per lane; one contiguous 64 KiB chunk per workgroup),
    # same stream, same bytes.
    stream_hip = {}
    for v, name in ((0, "word"), (1, "nontemporal"), (2, "four_words"), (3, "chunk_64KiB"), (4, "eight_words")):
        stream_hip[name] = round(state.stream_rate(v, 5)[0], 1)
    log(f"in-place streaming kernels over the state: {stream_hip} GB/s")
    del state
    ceiling = inplace_stream_rate(torch, (1 << n) * amp_bytes)
    log(f"in-place streaming rate of a plain torch multiply over the same bytes: {ceiling:.0f} GB/s")

    ms_per_step = gpu_ms / args.steps
    layers_per_s = args.depth / (ms_per_step / 1e3)
    launches = stats["passes"] * args.steps
    avg_launch_s = (gpu_ms / 1e3) / launches
    bytes_per_launch = 2 * (1 << n) * amp_bytes
    achieved = bytes_per_launch / avg_launch_s / 1e9

    out = {
        "metric": f"{n}q gate-layers/s",
        "value": round(layers_per_s, 3),
        "unit": "layers/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128" if args.precision == "complex128" else "c64",
        "data": f"synthetic: seeded random circuit (SmallRng seed {args.seed}), {args.depth} layers",
        "config": {
            "workload": f"random layered circuit, {n} qubits, depth {args.depth}, "
                        f"{'complex128' if amp_bytes == 16 else 'complex64'}, fused tile passes",
            "qubits": n, "depth": args.depth, "ops": stats["ops_in"], "passes_per_step": stats["passes"],
            "tile_qubits": tile_bits, "register_bits": reg_bits, "phases_per_step": stats["phases"],
            "pass_kernels": "compiled" if compiled else ("register" if reg_bits else "lds"),
            "kernels_run": ran,
            "compile_ms": round(compile_info["compile_ms"], 1) if compile_info else None,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "kernel": ("qf_pass_* (one compiled kernel per tile pass)" if compiled
                       else ("k_tile_reg" if reg_bits else "k_tile")),
            "bytes_per_launch": bytes_per_launch,
            "avg_launch_ms": round(avg_launch_s * 1e3, 4),
            "inplace_stream_GBs": round(ceiling, 1),
            "stream_kernels_GBs": stream_hip,
            "frac_of_stream_ceiling": round(achieved / max(max(stream_hip.values()), ceiling), 4),
        },
        "cpu_baseline": None,
    }
    out["roofline"]["traffic"] = measured_traffic(n, args, tile_bits, desc["passes"][0].get("low", 0), reg_bits,
                                                  stats["passes"], compiled)
    if args.c64_steps > 0 and args.precision == "complex128":
        out["complex64"] = complex64_section(be, torch, circuit, n, args.c64_steps, 1)
    if args.vqe_sets > 0:
        out["vqe_adjoint"] = vqe_bench(be, args.vqe_sets, args.vqe_shift_sets, not args.no_cpu)
    if args.small_steps > 0:
        c20 = random_layer_circuit(20, args.depth, seed=args.seed)
        r = circuit_section(be, torch, c20, 20, "complex128", args.small_steps, 2,
                            f"random layered circuit, 20 qubits, depth {args.depth}, complex128 (16 MiB: cache-resident)")
        r["layers_per_s"] = round(args.depth / (r["ms_per_circuit"] * 1e-3), 2)
        out["small_20q"] = r
    if args.qft_steps > 0:
        out["qft_like"] = qft_like_bench(be, torch, args.qft_steps)
    if args.pcie_qubits > 0:
        out["host_boundary"] = pcie_bench(be, args.pcie_qubits)
        log(f"host boundary: {out['host_boundary']}")
    if args.density_qubits > 0:
        out["density"] = density_bench(be, torch, args.density_qubits, args.density_layers, args.density_steps,
                                       not args.no_cpu)
    if args.traj_qubits > 0:
        out["trajectories"] = trajectory_bench(be, args.traj_qubits, args.traj_shots, 2, not args.no_cpu)
    if args.lanczos_sites > 0 or args.lanczos_csr_sites > 0:
        out["lanczos"] = lanczos_bench(be, args.lanczos_sites, args.lanczos_csr_sites, not args.no_cpu)
    if args.uneven_instances > 0:
        out["uneven_qaoa"] = uneven_bench(be, args.uneven_instances, 512, 0 if args.no_cpu else args.uneven_check)
    if not args.no_cpu:
        log("timing the CPU baseline (oracle restatement of the reference CPU path)")
        out["cpu_baseline"] = cpu_baseline(circuit, n, args.cpu_sample_ops)
    if args.fidelity_qubits > 0 and not args.no_cpu:
        out["fidelity_vs_oracle"] = fidelity_oracle_section(be, args.fidelity_qubits, args.seed)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

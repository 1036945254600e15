"""Benchmark: chi0 + W_c + Sigma_c per q-point on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--f64-steps K64] [--config C3]
                    [--mixed bf16x3|bf16f8|f32] [--no-cpu-baseline] [--also LIST] [--c4-eager]

A step is one q-point of the path (chi0(q) over all k and band pairs, W_c(q)
at nf+1 frequencies, Sigma_c(q) over all internal k) on synthetic inputs
already resident in device memory.  The headline workload is the
mixed-precision q-point (C3: nk=64, nb=320, 18+300 bands, nI=1200, nf=20),
timed in the mixed form (K steps after W warm-up steps) and in the plain
FP64 GPU form (K64 steps); `value` is the mixed form's seconds per q-point,
with its speedup over the FP64 form and the largest QP-energy difference
between the two (meV).  Prints one JSON line.  The kernel roofline uses HIP
events recorded by the library on its own stream around every launch of the
dominant kernel (the Sigma_c frequency contraction) during the timed steps.
Timed regions are bracketed by device synchronisation through the library
(ecalj_sync), which owns the process's only HIP runtime.  cpu_baseline times
the FP64 C oracle on a bounded sample of the same q-point (one k of chi0 and
of Sigma_c, one frequency of the W build, ...; 1 warm-up + median of 5 per
unit), scaled to a whole q-point.  Other workloads (C2, C4, C5, the real axis,
...) are reported under `also`.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"f64": 78.6, "f32": 157.3, "bf16x3": 2500.0, "bf16f8": 2500.0}  # TFLOP/s, MI355X dense

# HBM traffic per launch of the dominant kernel, from the calibrated FETCH_SIZE +
# WRITE_SIZE passes whose summary is committed (the counters cannot run inside
# this timed process); None where no pass was taken.
# HBM/fabric bytes per launch of the Sigma_c kernel from rocprofv3 FETCH_SIZE / WRITE_SIZE runs
# (calibrated on the copy kernel), per workload, form and the variant the tuner keeps
MEASURED_TRAFFIC = {
    ("C2", "bf16x3", "128x128x2/8w/xcd"): (49.7e9, "profiles/sigc_bf3_c2_mixed_hbm_traffic.md"),
    ("C2", "bf16x3", "128x128x2/8w"): (124.0e9, "profiles/sigc_bf3_c2_mixed_hbm_traffic.md"),
    ("C2", "bf16f8", None): (48.8e9, "profiles/sigc_f8_c2_hbm_traffic.md"),  # one kernel, XCD-aware order
    ("C3", "bf16x3", "3m/128x128/xcd"): (50.83e9, "profiles/sigc_3m_c3_hbm_traffic.md"),
    ("C3", "bf16x3", "3m/128x128/w8/xcd"): (56.37e9, "profiles/sigc_3m_w8_c3_hbm_traffic.md"),
}


def chosen_variant(tuning):
    """The Sigma_c variant the tuner kept (marked * in its report), or None."""
    for line in (tuning or "").splitlines():
        if line.startswith("sigc_bf3 ") and "*" in line:
            toks = line.split("]: ", 1)[-1].split()  # name time "ms" ["*"] ...
            for i, t in enumerate(toks):
                if t == "*" and i >= 3:
                    return toks[i - 3]
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(qp, reps=5):
    """FP64 oracle on the host cores, timed on a bounded sample of the q-point
    and scaled to one q-point.  Each unit of the sample runs once as a warm-up
    and then `reps` times; the median counts.  The oracle's GEMMs and LU
    inverses go through the LP64 zgemm / zgetrf / zgetri scipy ships
    (OpenBLAS), where the reference calls MKL's zgemm3m / zgetrf / zgetri
    (S/m_blas.f90:225, S/m_lapack.f90:21-47); the loops around them are the
    oracle's own (OpenMP).  Units: zmel + x0gemm for one k of chi0 (x nk);
    the Hilbert transform on an eighth of rcxq's columns (x 8); the W build
    at one frequency (x nf+1); Sigma_c's zmel + weights for one internal k
    (x nk); the Sigma_c contraction of one k at one and at two frequencies
    (the difference is one frequency: x nk (nf+1); the rest, the final
    product, x nk)."""
    import oracle as orc
    from ecalj_amd import synth
    from oracle import pipeline

    orc.build()
    cfg = qp.cfg
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    try:
        from threadpoolctl import threadpool_limits
        limiter = threadpool_limits(limits=threads, user_api="blas")
    except ImportError:
        limiter = None
    blas = orc.use_blas(True)

    def med(f):
        f()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            f()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts))

    ni, nk, nw1 = cfg.ni, cfg.nk, cfg.nf + 1
    fr = synth.frhis(qp)
    nwhis = len(fr) - 1
    pw = synth.pair_weights(qp, fr)
    ix = orc.x0kf_init([(pw.n1b, pw.n2b, pw.ihw, pw.nhw, pw.jhw)], pw.whw, nband=cfg.nbt, nwhis=nwhis)
    no, nu = cfg.no, cfg.nu
    one = [1.0] * cfg.natom
    t = {}
    z = orc.zmel(cfg.nv_list, cfg.mdim_list, one, qp.ppb, qp.c_k[:, :no, 0], qp.c_kq[:, no:no + nu, 0], conjg=True)
    t["zmel_chi0_k"] = med(lambda: orc.zmel(cfg.nv_list, cfg.mdim_list, one, qp.ppb, qp.c_k[:, :no, 0],
                                            qp.c_kq[:, no:no + nu, 0], conjg=True))
    rc = np.zeros((ni, ni, nwhis + 1), np.complex128, order="F")
    t["x0gemm_k"] = med(lambda: orc.x0gemm(rc, ni, 1, ni, nwhis, 1, int(ix.nkmin[0]), int(ix.nkmax[0]), z,
                                           int(ix.icounkmin[0]), int(ix.icounkmax[0]), ix.iwini, ix.iwend, ix.itc,
                                           ix.itpc, ix.jpmc, ix.icouini, ix.whwc))
    ncol = max(1, ni // 8)
    slab = np.asfortranarray(rc[:, :ncol, :])
    del rc
    _, _, _, _, nu_ = pipeline.frequencies(cfg)
    im = orc.imatt(nu_, fr)
    t["hilbert_eighth"] = med(lambda: orc.dpsion_imag(slab, fr, im))
    del slab
    # the W build's cost does not depend on the values: one Hermitian negative definite chi
    xr = np.random.default_rng(0).standard_normal((ni, 64)) * (0.1 / ni)
    chi1 = np.asfortranarray(-(xr @ xr.T)[:, :, None].astype(np.complex128))
    t["w_freq"] = med(lambda: orc.wc_build(chi1, qp.vcousq, qp.zcousq, nspin=cfg.nspin))
    wc1 = orc.wc_build(chi1, qp.vcousq, qp.zcousq, nspin=cfg.nspin)
    freqx, _, wx, expa, _ = pipeline.frequencies(cfg)
    omega = qp.e_q[:cfg.nq] / synth.RYDBERG_EV
    nbt = cfg.nbt

    def sig_prep():
        zs = orc.zmel(cfg.nv_list, cfg.mdim_list, one, qp.ppb, qp.c_qmk[:, :nbt, 0], qp.c_q[:, :cfg.nq])
        g = orc.wgtim(omega, qp.e_qmk[:nbt, 0] / synth.RYDBERG_EV, cfg.esmr, qp.wk[0], freqx, wx, expa, ua=cfg.ua)
        return zs, g

    t["zmel_wgt_sigma_k"] = med(sig_prep)
    zs, g = sig_prep()
    w2 = np.asfortranarray(np.concatenate([wc1, wc1], axis=2))
    t["sigc_k_1freq"] = med(lambda: orc.sigc_imag(zs, wc1, g[:1]))
    t["sigc_k_2freq"] = med(lambda: orc.sigc_imag(zs, w2, g[:2]))
    tf = max(t["sigc_k_2freq"] - t["sigc_k_1freq"], 0.0)
    tfin = max(t["sigc_k_1freq"] - tf, 0.0)
    per_q = ((t["zmel_chi0_k"] + t["x0gemm_k"]) * nk + t["hilbert_eighth"] * (ni / ncol) + t["w_freq"] * nw1
             + t["zmel_wgt_sigma_k"] * nk + (tf * nw1 + tfin) * nk)
    if limiter is not None:
        limiter.restore_original_limits()
    measured = sum(t.values()) * (reps + 1)
    sample = (f"{cfg.name}: FP64 oracle ({'OpenBLAS zgemm/zgetrf/zgetri' if blas else 'own loops'} + OpenMP), "
              f"units timed 1 warm-up + median of {reps}: zmel+x0gemm of 1 of {nk} k, Hilbert on {ncol} of {ni} "
              f"columns, W build at 1 of {nw1} frequencies, Sigma_c zmel+weights of 1 of {nk} k, Sigma_c "
              f"contraction of 1 k at 1 and 2 of {nw1} frequencies; scaled to one q-point "
              f"({measured:.1f} s of CPU work)")
    return {"value": round(per_q, 3), "unit": "s/q-point", "cores": threads, "kind": "port", "sample": sample,
            "units_s": {k: round(v, 4) for k, v in t.items()}}


def run_form(qp, precision, steps, warmup, tuning=0):
    """Time `steps` q-point steps of one precision form; returns timings and QP energies.
    tuning > 0 forces that Sigma_c kernel variant (profiling one variant)."""
    from ecalj_amd.qsgw import QPointGPU

    g = QPointGPU(qp, precision)
    if tuning:
        g.h.set_tuning(tuning)
    log(f"bench[{precision}]: setup run (W_c at internal k from the first W_c(q))")
    g.step()
    g.make_wk()
    g.h.sync()
    tuning = g.h.tuning_report().strip()
    if tuning:
        log(f"bench[{precision}]: tuner: {tuning}")
    for i in range(warmup):
        g.step()
        g.h.sync()
        log(f"bench[{precision}]: warmup {i + 1}/{warmup} done")
    g.h.kernel_timing(True)
    g.h.sync()
    t0 = time.perf_counter()
    for i in range(steps):
        g.step()
    g.h.sync()
    dt = time.perf_counter() - t0
    ktimes = {k: g.h.kernel_time(k) for k in ("sigc", "chi0", "zmel", "wbuild", "hilbert")}
    g.h.kernel_timing(False)
    _, _, _, eqp = g.results()
    log(f"bench[{precision}]: {steps} steps in {dt:.3f} s -> {dt / steps * 1e3:.1f} ms per q-point")
    info = {"s_per_q": dt / steps, "ktimes": ktimes, "eqp": eqp, "nwhis": g.nwhis, "tuning": tuning,
            "npb": int(sum(g.ix.iwend - g.ix.iwini + 1))}
    g.close()
    del g
    return info


def under_profiler() -> bool:
    """rocprofv3 exports its options as ROCPROF_* to the program it runs; the
    library refuses HIP-graph capture there, so the C4 legs run eagerly."""
    return any(k in os.environ for k in ("ROCPROF_OUTPUT_PATH", "ROCPROF_KERNEL_TRACE", "ROCPROF_STATS"))


def bench_c4(mixed, nq=16, log_=None, graph=True):
    """Batched uneven workload, one GPU's share: nq q-points of 27 k-points
    with per-k band windows and per-q nI; FP64 form eager, mixed form as a
    captured HIP graph per q-point (replayed, as the QSGW iterations do), or
    eagerly only (graph=False, for runs under a profiler)."""
    from ecalj_amd import synth
    from ecalj_amd.batched import UnevenQPointGPU, flops_uneven

    graph = graph and not under_profiler()
    rows = []
    t64 = tmx = tmx_eager = fl = 0.0
    err = 0.0
    for c in synth.c4_qpoints(nq):
        qp = synth.make_uneven_qpoint(c, no_range=(8, 24), nu_range=(150, 400))
        res = {}
        for prec in ("f64", mixed):
            g = UnevenQPointGPU(qp, prec)
            g.step()
            g.make_wk()
            g.step()                      # warm-up at final shapes (tuner settles)
            g.h.sync()
            t0 = time.perf_counter()
            g.step()
            g.h.sync()
            te = time.perf_counter() - t0
            tg = None
            if prec != "f64" and not graph:
                tg = te
            elif prec != "f64":
                gr = g.capture_step()
                gr.launch()
                g.h.sync()
                t0 = time.perf_counter()
                gr.launch()
                gr.launch()
                g.h.sync()
                tg = (time.perf_counter() - t0) / 2
                gr.close()
            res[prec] = (te, tg, g.results()[3], g.ix)
            g.h.close()
        npb = int(sum(res["f64"][3].iwend - res["f64"][3].iwini + 1))
        f = flops_uneven(qp, npb)
        fl += f
        t64 += res["f64"][0]
        tmx += res[mixed][1]
        tmx_eager += res[mixed][0]
        e = float(np.abs(res[mixed][2] - res["f64"][2]).max() * 1e3)
        err = max(err, e)
        rows.append({"ni": c.ni, "f64_s": round(res["f64"][0], 4), "mixed_graph_s": round(res[mixed][1], 4),
                     "mixed_eager_s": round(res[mixed][0], 4), "qp_err_meV": round(e, 5)})
        if log_:
            log_(f"bench[C4]: q nI={c.ni}: f64 {res['f64'][0]:.3f} s, {mixed} graph {res[mixed][1]:.3f} s "
                 f"(eager {res[mixed][0]:.3f} s), dQP {e:.4f} meV")
    return {
        "workload": f"C4 one-GPU share: {nq} q-points, nI ~ U[800,1600] per q, 27 k-points with per-k windows "
                    f"U{{8..24}} occ x U{{150..400}} unocc, nf=20, ntq=16",
        "mixed_timing": "HIP-graph replay" if graph else "eager (no graph capture)",
        "s_per_q": round(tmx / nq, 6), "fp64_form_s_per_q": round(t64 / nq, 6),
        "mixed_eager_s_per_q": round(tmx_eager / nq, 6),
        "speedup_vs_fp64_form": round(t64 / tmx, 3), "qp_err_meV_vs_fp64_form": round(err, 5),
        "tflops": round(fl / tmx / 1e12, 3), "per_q": rows,
    }


def bench_c5(log_=None):
    """Learned chi0 tail: train on 4096 (truncated, full) pairs from 4 seeded
    q-points, judge on a held-out one; time the learned form (chi0 truncated
    at nu_t + tail) against the full band sum, both in the mixed form."""
    from ecalj_amd import learned_tail as lt
    from ecalj_amd import synth
    from ecalj_amd.qsgw import QPointGPU

    cfg = synth.CONFIGS["C5"]
    m = lt.fit(lt.seeded_qpoints(cfg, 4), cfg.nu_chi0, npairs=4096, epochs=300, log=log_)
    held = lt.seeded_qpoints(cfg, 1, first=4)
    j = lt.judge(m, held, tol_mev=1.0, log=log_)
    qp = held[0]
    times = {}
    chi0_t = {}
    for name, kw, q in (("full_fp64", {}, synth.dataclasses.replace(qp, cfg=cfg.scaled(nu_chi0=0))),
                        ("full_mixed", {}, synth.dataclasses.replace(qp, cfg=cfg.scaled(nu_chi0=0))),
                        ("learned_mixed", {"tail_model": m}, qp)):
        g = QPointGPU(q, "f64" if name == "full_fp64" else "bf16x3", **kw)
        g.step()
        g.make_wk()
        g.step()
        g.h.sync()
        t0 = time.perf_counter()
        g.step()
        g.h.sync()
        times[name] = time.perf_counter() - t0
        t0 = time.perf_counter()
        g.chi0()  # chi0 alone (the learned form: truncated sum + tail)
        g.h.sync()
        chi0_t[name] = time.perf_counter() - t0
        g.h.close()
    bound = lt.tail_bound(qp, cfg.nu_chi0, ndraw=7, log=log_)
    r = j["rows"][0]
    return {
        "workload": f"C5: nI={cfg.ni}, nf={cfg.nf}, nk={cfg.nk}, {cfg.no} occ, chi0 truncated at {cfg.nu_chi0} of "
                    f"{cfg.nu} unocc; MLP {m.dims} trained on {m.info['npairs']} pairs by the library's kernels",
        "train_s": round(m.info["t_train_s"], 3), "final_loss": m.info["losses"][-1],
        "qp_err_meV_truncated": round(r["qp_err_truncated_meV"], 3),
        "qp_err_meV_learned": round(r["qp_err_learned_meV"], 3),
        "chi_diag_err_truncated": round(r["chi_diag_err_truncated"], 5),
        "chi_diag_err_learned": round(r["chi_diag_err_learned"], 5),
        "judged_error_meV": round(j["judged_error_meV"], 3), "accepted": j["accepted"],
        "tail_bound": {k: (round(v, 3) if isinstance(v, float) else v) for k, v in bound.items()},
        "form_used": "learned tail" if j["accepted"] else "full band sum (fallback)",
        "s_per_q": {k: round(v, 5) for k, v in times.items()},
        "chi0_s": {k: round(v, 5) for k, v in chi0_t.items()},
        "chi0_share_of_step_full_mixed": round(chi0_t["full_mixed"] / times["full_mixed"], 3),
    }


def bench_real_axis(mixed, name="C2", log_=None):
    """The q-point with the real-axis chain added: chi(omega) on freq_r from the
    same rcxq, W_c(omega) (WVRllwR, pivoted inversion) and the Sigma_c pole
    term in the same contraction.  Both forms, stage times with a sync between
    stages, one warm-up step."""
    from ecalj_amd import synth
    from ecalj_amd.qsgw import QPointGPU

    c = synth.CONFIGS[name]
    qp = synth.make_qpoint(c)
    res = {}
    for prec in ("f64", mixed):
        g = QPointGPU(qp, prec, real_axis=True)
        g.step()
        g.make_wk()
        g.step()
        g.h.sync()
        st = {}
        t0 = time.perf_counter()
        for stage in ("chi0", "wbuild", "sigc"):
            t1 = time.perf_counter()
            getattr(g, stage)()
            g.h.sync()
            st[stage] = round(time.perf_counter() - t1, 5)
        tot = time.perf_counter() - t0
        res[prec] = (tot, st, g.results()[3], g.nwr)
        if log_:
            log_(f"bench[{name} real axis]: {prec} {tot:.4f} s {st}")
        g.h.close()
    e = float(np.abs(res[mixed][2] - res["f64"][2]).max() * 1e3)
    return {
        "workload": f"{name} q-point plus the real-axis chain: chi(omega) and W_c(omega) at nw+1={res['f64'][3] + 1} "
                    f"real frequencies (freq_r from getwemax/getfreq), pole term of Sigma_c for all {c.nk} k",
        "s_per_q": round(res[mixed][0], 5), "fp64_form_s_per_q": round(res["f64"][0], 5),
        "speedup_vs_fp64_form": round(res["f64"][0] / res[mixed][0], 3),
        "qp_err_meV_vs_fp64_form": round(e, 5),
        "stage_s": {"f64": res["f64"][1], mixed: res[mixed][1]},
    }


def bench_tetwt(log_=None):
    """Tetrahedron histogram weights (gettetwt + tetwt5x_dtet4): the C2-like
    mesh (4x4x4 k, 9 occupied + 150 unoccupied bands) on the GPU and in the
    FP64 oracle (one host thread), and the C4-like mesh (6x6x6 k, 24 + 400
    bands) on the GPU.  Smooth synthetic bands, two crossing E_F."""
    from ecalj_amd import Handle, synth

    out = {}
    h = Handle(0, "f64")
    try:
        for name, n, nocc, nun, cpu in (("C2-like", (4, 4, 4), 9, 150, True), ("C4-like", (6, 6, 6), 24, 400, False)):
            m = synth.tetra_mesh(*n)
            ek1, ek2, ef = synth.tetra_bands(m, nocc, nun)
            fr = synth.tetra_frhis(ek1, ek2)
            h.tetwt(m, ek1, ek2, ef, fr)
            ts = []
            for _ in range(3):
                t0 = time.perf_counter()
                r = h.tetwt(m, ek1, ek2, ef, fr)
                ts.append(time.perf_counter() - t0)
            row = {"nqbz": m.nqbz, "ntetf": int(m.idtetf.shape[1]), "nband": nocc + nun,
                   "pairs": int(r["nbnb"].sum()), "nhwtot": int(r["nhwtot"]), "gpu_s": round(float(np.median(ts)), 5)}
            if cpu:
                import oracle as orc
                orc.build()
                t0 = time.perf_counter()
                o = orc.tetwt(m, ek1, ek2, ef, fr)
                row["cpu_oracle_s"] = round(time.perf_counter() - t0, 4)
                row["cpu_threads"] = 1
                row["index_maps_equal"] = bool(all(np.array_equal(r[k], o[k]) for k in ("n1b", "n2b", "ihw", "nhw", "jhw")))
            if log_:
                log_(f"bench[tetwt {name}]: {row}")
            out[name] = row
    finally:
        h.close()
    out["note"] = ("host-pointer call: includes the uploads, two small host round trips between the passes and "
                   "the copy-out of the lists and whw")
    return out


def bench_ipw(mixed, log_=None):
    """Interstitial plane-wave rows of zmel for one k at GaAs-like sizes: 150
    IPW product functions, 180 plane waves at q and at q-k; the Sigma_c shape
    (159 middle x 18 end states: G1 first) and the chi0 shape (9 occupied x
    150 unoccupied: G2 first).  Device time per call (inputs resident), both
    forms; the FP64 oracle on one host thread for comparison."""
    import oracle as orc
    from ecalj_amd import Handle, synth

    orc.build()
    out = {}
    for name, nm, ntp0 in (("sigma_shape", 159, 18), ("chi0_shape", 9, 150)):
        c = synth.ipw_case(7, 150, 180, 180, nm, ntp0, gcut2=40)
        row = {"ngc": 150, "ngp1": 180, "ngp2": 180, "nm": nm, "ntp0": ntp0, "ngcgp": int(c.nvgcgp2.shape[1])}
        ref = None
        for prec in ("f64", mixed):
            h = Handle(0, prec)
            try:
                h.set_ppovl(c)
                st = np.complex128 if prec == "f64" else np.complex64
                dq, dk = h.to_dev(c.geigq, np.complex128), h.to_dev(c.dgeigqk, np.complex128)
                dz = h.array((150, nm, ntp0), st)
                h.zmel_ipw_dev(c, dq, dk, dz, 150)
                h.sync()
                t0 = time.perf_counter()
                for _ in range(10):
                    h.zmel_ipw_dev(c, dq, dk, dz, 150)
                h.sync()
                row[f"{prec}_ms"] = round((time.perf_counter() - t0) / 10 * 1e3, 4)
                got = dz.download().astype(np.complex128)
            finally:
                h.close()
            if ref is None:
                t0 = time.perf_counter()
                ref = orc.zmel_ipw(c)
                row["cpu_oracle_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
            row[f"{prec}_relerr"] = float(np.abs(got - ref).max() / np.abs(ref).max())
        if log_:
            log_(f"bench[IPW {name}]: {row}")
        out[name] = row
    return out


def bench_rotwave(log_=None):
    """Eigenfunctions at 66 points (3 irreducible + 21 rotations of each) from
    the 3 irreducible ones (readcphi + readgeig): 180 bands, 34 muffin-tin rows
    with l <= 3 on two atoms, plane-wave lists of |q+G| < 4.6 (up to 414 G).  Device
    time per batch (inputs resident, host index maps included), bytes moved
    (read + write of the rotated blocks), and the FP64 oracle on one target."""
    import oracle as orc
    from ecalj_amd import Handle, synth

    orc.build()
    c = synth.rotation_case(seed=4, nqi=3, ntarget=21, qcut=4.6)
    nk, nband = c.qtt.shape[1], 180
    rng = np.random.default_rng(1)
    cphi = (rng.standard_normal((c.ndima, nband, c.nqi)) + 1j * rng.standard_normal((c.ndima, nband, c.nqi)))
    geig = (rng.standard_normal((c.ngpmx, nband, c.nqi)) + 1j * rng.standard_normal((c.ngpmx, nband, c.nqi)))
    iq = np.arange(1, nk + 1)
    h = Handle(0, "f64")
    row = {"targets": nk, "nband": nband, "ndima": c.ndima, "ngpmx": c.ngpmx}
    try:
        h.set_hamindex(c)
        h.set_mnla(c.ibas_tbl, c.l_tbl, c.k_tbl, c.offset_tbl)
        dc, dg = h.to_dev(np.asfortranarray(cphi)), h.to_dev(np.asfortranarray(geig))
        oc, og = h.array((c.ndima, nband, nk)), h.array((c.ngpmx, nband, nk))
        for _ in range(2):
            h.readcphi_dev(iq, 1, nband, dc, c.ndima, oc)
            h.readgeig_dev(iq, 1, nband, dg, c.ngpmx, og)
        h.sync()
        reps = 20
        t0 = time.perf_counter()
        for _ in range(reps):
            h.readcphi_dev(iq, 1, nband, dc, c.ndima, oc)
            h.readgeig_dev(iq, 1, nband, dg, c.ngpmx, og)
        h.sync()
        dt = (time.perf_counter() - t0) / reps
        moved = 2 * 16 * nband * nk * (c.ndima + c.ngpmx) + 16 * nband * nk * c.ngpmx  # + zero fill of geig
        row["ms_per_batch"] = round(dt * 1e3, 4)
        row["bytes_per_batch"] = moved
        row["GB_per_s"] = round(moved / dt / 1e9, 1)
        got_c, got_g = oc.download(), og.download()
    finally:
        h.close()
    t0 = time.perf_counter()
    wc = orc.rotmto(c, nk, cphi[:, :, c.iqimap[nk - 1] - 1])
    wg = orc.rotipw(c, nk, np.asfortranarray(geig[:c.ngp[c.iqmap[nk - 1] - 1], :, c.iqimap[nk - 1] - 1]))
    row["cpu_oracle_ms_per_target"] = round((time.perf_counter() - t0) * 1e3, 3)
    row["relerr_vs_oracle"] = float(max(np.abs(got_c[:, :, -1] - wc).max() / np.abs(wc).max(),
                                        np.abs(got_g[:wg.shape[0], :, -1] - wg).max() / np.abs(wg).max()))
    if log_:
        log_(f"bench[ROT]: {row}")
    return row


def bench_gamma(mixed, log_=None):
    """The Gamma point of the W build at C2 sizes (nI = 600): W(q=0) on the 16
    imaginary and the nw+1 real frequencies of the C2 real mesh (mu = 1 left
    out of the inversion), llw = 1/epsinv(1,1) at six offset points on both
    meshes (the wing at omega = 0), getw0 for both, and modifyWV0 with the
    E -> M transform.  Both forms (the mixed one with the M <-> E transforms on
    its core around the FP64 inverse); inputs resident; device time per Gamma
    point.  The FP64 oracle on a one-frequency sample of each inversion kind,
    scaled to the same work."""
    import oracle as orc
    from ecalj_amd import Handle, synth

    orc.build()
    c = synth.CONFIGS["C2"]
    qp = synth.make_qpoint(c)
    nwr = len(synth.real_mesh(qp, synth.frhis(qp)))
    n, niw = c.ni, c.nf
    rng = np.random.default_rng(5)
    gc = synth.gamma_cell(lmxax=4)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    Q = np.asfortranarray(Q)
    vc = rng.uniform(0.5, 40.0, n)
    vcou1 = 4 * np.pi / np.sum(gc.q0i[:, 0] ** 2)
    vc[0] = vcou1 * 0.999

    def chi(nw, real):
        x = (rng.standard_normal((n, 40)) + 1j * rng.standard_normal((n, 40))) * 0.004
        base = -(x @ x.conj().T)
        out = np.empty((n, n, nw), complex, order="F")
        for i in range(nw):
            out[:, :, i] = base / (1 + 0.1 * i)
            if real:
                out[:, :, i] += 1j * 0.3 * base / (1 + 0.05 * i)
        out[0, :, :] /= np.sqrt(vcou1)
        out[:, 0, :] /= np.sqrt(vcou1)
        return out

    ci, cr = chi(niw, False), chi(nwr, True)
    row = {"ngb": n, "niw": niw, "nw_real": nwr, "nq0i": 6}
    w0_forms = {}
    for prec in ("f64", mixed):
        h = Handle(0, prec)
        try:
            h.set_coulomb(vc, Q)
            di, dr = h.to_dev(ci), h.to_dev(cr)
            wi, wr = h.array((n, n, niw)), h.array((n, n, nwr))
            li, lr = h.array((niw,)), h.array((nwr,))
            wing = h.array((n - 1,))
            llw_i = np.zeros((niw, 6), complex)
            llw_r = np.zeros((nwr, 6), complex)

            def step():
                for k in range(6):  # one offset point each; same chi for the timing
                    h.eps_head_dev(di, n, niw, vcou1, li, real_axis=False)
                    h.eps_head_dev(dr, n, nwr, vcou1, lr, real_axis=True, wing_iw=0, wing=wing)
                h.sync()
                llw_i[:] = li.download()[:, None]
                llw_r[:] = lr.download()[:, None]
                w0i, _ = ecalj_getw0(1, llw_i, gc)
                w0r, _ = ecalj_getw0(0, llw_r, gc)
                h.wc_build_gamma_dev(di, n, niw, wi, real_axis=False)
                h.wc_build_gamma_dev(dr, n, nwr, wr, real_axis=True)
                h.wv0_patch_dev(wi, n, niw, w0i)
                h.wv0_patch_dev(wr, n, nwr, w0r)
                h.sync()
                return w0i, w0r

            step()
            t0 = time.perf_counter()
            w0_forms[prec] = step()
            row[f"{prec}_s_per_gamma_point"] = round(time.perf_counter() - t0, 5)
            for name, fn in (("eps_head_real_ms", lambda: h.eps_head_dev(dr, n, nwr, vcou1, lr, real_axis=True)),
                             ("wc_build_gamma_real_ms",
                              lambda: h.wc_build_gamma_dev(dr, n, nwr, wr, real_axis=True))):
                h.sync()
                t0 = time.perf_counter()
                fn()
                h.sync()
                row[f"{prec}_{name}"] = round((time.perf_counter() - t0) * 1e3, 3)
        finally:
            h.close()
    a, b = w0_forms["f64"], w0_forms[mixed]
    row["w0_relerr_mixed_vs_f64"] = float(max(np.abs(a[0] - b[0]).max() / np.abs(a[0]).max(),
                                              np.abs(a[1] - b[1]).max() / np.abs(a[1]).max()))
    t0 = time.perf_counter()
    orc.eps_head(ci[:, :, :1], np.sqrt(vc), Q, vcou1)
    t1 = time.perf_counter()
    orc.wc_build_gamma(ci[:, :, :1], np.sqrt(vc), Q)
    t2 = time.perf_counter()
    row["cpu_oracle_s_scaled"] = round((t1 - t0) * 6 * (niw + nwr) + (t2 - t1) * (niw + nwr), 2)
    row["cpu_sample"] = "one eps_head and one wc_build_gamma frequency at nI=600 (oracle, host cores), scaled"
    if log_:
        log_(f"bench[GAM]: {row}")
    return row


def ecalj_getw0(ii, llw, gc):
    import ecalj_amd

    return ecalj_amd.getw0(ii, llw, gc.dmlx, gc.epinvq0i, gc.wklm, gc.wbz, gc.lmxax, gc.epinv)


def roofline(fl, info, precision, steps, config="C2"):
    ms, nl = info["ktimes"]["sigc"]
    launches_per_step = max(1, nl // steps)
    flop_launch = fl["sigc"] / launches_per_step
    avg_s = ms / 1e3 / max(nl, 1)
    achieved = flop_launch / avg_s / 1e12
    peak = PEAK[precision]
    var = chosen_variant(info.get("tuning"))
    kname = {"bf16x3": "sig_bf3_kernel (bf16x3 Sigma_c: per-frequency W_c(iv)^H zmel, weighted sum, all internal k)",
             "bf16f8": "sig_f8_kernel (Sigma_c: bf16 hi*hi + scaled-fp8 cross terms, per-frequency W_c(iv)^H zmel, "
                       "weighted sum, all internal k)"}.get(
        precision, "sig_freq_gemm_kernel (Sigma_c: W_c(iv)^H x frequency-weighted zmel, all internal k)")
    if precision == "bf16x3" and var and var.startswith("3m32"):
        kname = ("sig3m32_kernel (bf16x3 Sigma_c, frequency-inner, three real products per complex product on the "
                 "32x32x16 MFMA: W_c(iv)^H x frequency-weighted zmel, all internal k)")
    elif precision == "bf16x3" and var and var.startswith("3m"):
        kname = ("sig3m_kernel (bf16x3 Sigma_c, frequency-inner, three real products per complex product: "
                 "W_c(iv)^H x frequency-weighted zmel, all internal k)")
    elif precision == "bf16x3" and var and var.startswith("vin4"):
        kname = ("sigv4_kernel (bf16x3 Sigma_c, frequency-inner, four real products per complex product: "
                 "W_c(iv)^H x frequency-weighted zmel, all internal k)")
    tr = MEASURED_TRAFFIC.get((config, precision, var))
    r = {"kernel": kname,
         "bound": "mfma", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
         "frac": round(achieved / peak, 4), "traffic": tr[0] if tr else None,
         "flop_per_launch": flop_launch, "avg_launch_ms": round(avg_s * 1e3, 4)}
    if tr:
        r["traffic_unit"] = "bytes per launch (HBM/fabric reads + writes, calibrated counters)"
        r["traffic_source"] = tr[1]
    if precision == "bf16x3":
        r["variant"] = var
        if var and var.startswith("3m"):
            r["note"] = ("achieved counts algorithmic complex flops (8 per complex MAC); the three-product form "
                         "costs 9 bf16 MFMA products per 4 real products (3 complex products x 3 split terms), "
                         "2.25 MFMA flops per algorithmic flop, so the attainable algorithmic rate is peak/2.25")
        else:
            r["note"] = ("achieved counts algorithmic complex flops (8 per complex MAC); each costs 3 bf16 MFMA "
                         "products, so the attainable algorithmic rate is peak/3")
    if precision == "bf16f8":
        r["note"] = ("achieved counts algorithmic complex flops (8 per complex MAC); each costs one bf16 MFMA "
                     "product and two fp8 products at twice the bf16 rate, so the attainable algorithmic rate "
                     "is peak/2")
    return r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--f64-steps", type=int, default=3,
                    help="timed steps of the FP64 GPU form the speedup is taken against (C3: ~4 s each)")
    ap.add_argument("--config", default="C3")
    ap.add_argument("--mixed", default="bf16x3", choices=["bf16x3", "bf16f8", "f32"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--c4-eager", action="store_true",
                    help="C4 legs: time the mixed form eagerly only (no HIP-graph capture/replay)")
    ap.add_argument("--also", default="C2,F8,C2F8,C4,C4F8,C5,C2R,TW,IPW,GAM,ROT",
                    help="comma-separated further configs timed in both forms (2 steps each) and reported "
                         "under 'also'; 'none' to skip")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run the library's plain copy kernel on MIB MiB (HBM-counter calibration)")
    ap.add_argument("--also-only", action="store_true",
                    help="run only the --also legs (prints their JSON; for profiling them)")
    ap.add_argument("--only", choices=["f64", "mixed"], default=None,
                    help="profile one form only (prints a short line instead of the contract JSON)")
    ap.add_argument("--tuning", type=int, default=0, metavar="K",
                    help="with --only: force Sigma_c kernel variant K (1-based, as ecalj_set_tuning)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the q-points of a QSGW "
              "iteration are independent and would shard one per GPU). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from ecalj_amd import synth
    from ecalj_amd.qsgw import flops_per_q

    if args.copy_calibration:
        import ecalj_amd

        hc = ecalj_amd.Handle(0)
        nbytes = args.copy_calibration << 20
        a, b = hc.array(nbytes, np.uint8), hc.array(nbytes, np.uint8)
        a.zero()
        for _ in range(3):
            hc.dev_copy(b, a)
        hc.sync()
        log(f"bench: copy kernel ran 3x on {args.copy_calibration} MiB (reads and writes {nbytes} bytes each)")
        hc.close()
    if args.also_only:
        res = {}
        for name in [c for c in args.also.split(",") if c and c.lower() != "none"]:
            if name == "TW":
                res[name] = bench_tetwt(log_=log)
            elif name == "IPW":
                res[name] = bench_ipw(args.mixed, log_=log)
            elif name == "GAM":
                res[name] = bench_gamma(args.mixed, log_=log)
            elif name == "ROT":
                res[name] = bench_rotwave(log_=log)
            elif name in ("C4", "C4F8"):
                res[name] = bench_c4(args.mixed if name == "C4" else "bf16f8", log_=log, graph=not args.c4_eager)
            elif name == "C5":
                res[name] = bench_c5(log_=log)
            elif name.endswith("R") and name[:-1] in synth.CONFIGS:
                res[name] = bench_real_axis(args.mixed, name[:-1], log_=log)
            else:
                raise SystemExit(f"--also-only: {name} is not a separate leg")
        print(json.dumps(res), flush=True)
        return
    cfg = synth.CONFIGS[args.config]
    log(f"bench: generating {cfg.name} inputs")
    qp = synth.make_qpoint(cfg)
    if args.only:
        info = run_form(qp, "f64" if args.only == "f64" else args.mixed, args.steps, args.warmup, args.tuning)
        print(json.dumps({"form": args.only, "s_per_q": info["s_per_q"],
                          "kernel_ms_per_step": {k: v[0] / args.steps for k, v in info["ktimes"].items()}}))
        return
    f64 = run_form(qp, "f64", args.f64_steps, min(args.warmup, 1))
    mix = run_form(qp, args.mixed, args.steps, args.warmup)
    fl = flops_per_q(cfg, nwhis_pairs_bins=mix["npb"])
    ftot = sum(fl.values())
    qp_err_mev = float(np.abs(mix["eqp"] - f64["eqp"]).max() * 1e3)
    s_per_q = mix["s_per_q"]
    out = {
        "metric": "chi0+Sigma_c s/q-point & TFLOP/s, mixed vs FP64 GPU form at 1 meV QP error",
        "value": round(s_per_q, 6),
        "unit": "s/q-point",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(s_per_q * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"bf16x3": "bf16x3 split products, FP32 accumulation, complex64 storage; FP64 epsilon and its inverse",
                  "bf16f8": "bf16x3 split products; Sigma_c as bf16 hi*hi + scaled fp8 (E4M3) cross terms; FP32 "
                            "accumulation, complex64 storage; FP64 epsilon and its inverse",
                  "f32": "complex64, FP32 MFMA; FP64 epsilon and its inverse"}[args.mixed],
        "data": "synthetic (seeded; see ecalj_amd/synth.py)",
        "config": {"workload": f"{cfg.name} q-point: nk={cfg.nk}, nb={cfg.nb}, occ={cfg.no}, unocc={cfg.nu}, "
                               f"nI={cfg.ni}, nf={cfg.nf}(+nu=0), ntq={cfg.nq}, nwhis={mix['nwhis']}",
                   "form": args.mixed},
        "tflops": round(ftot / s_per_q / 1e12, 3),
        "fp64_form": {"s_per_q": round(f64["s_per_q"], 6), "tflops": round(ftot / f64["s_per_q"] / 1e12, 3),
                      "steps": args.f64_steps, "warmup": min(args.warmup, 1),
                      "kernel_ms_per_step": {k: round(v[0] / args.f64_steps, 3) for k, v in f64["ktimes"].items()},
                      "roofline": roofline(fl, f64, "f64", args.f64_steps, cfg.name)},
        "speedup_vs_fp64_form": round(f64["s_per_q"] / s_per_q, 3),
        "qp_err_meV_vs_fp64_form": round(qp_err_mev, 5),
        "flops_per_q": ftot,
        "kernel_ms_per_step": {k: round(v[0] / args.steps, 3) for k, v in mix["ktimes"].items()},
        "roofline": roofline(fl, mix, args.mixed, args.steps, cfg.name),
        "tuning": mix["tuning"],
    }
    if not args.no_cpu_baseline:
        log("bench: timing the CPU oracle on a bounded sample")
        try:
            out["cpu_baseline"] = cpu_baseline(qp)
        except Exception as e:  # report, never hide
            out["cpu_baseline"] = {"value": None, "error": repr(e)}
    also = [c for c in args.also.split(",") if c and c.lower() != "none" and c != args.config]
    if also:
        out["also"] = {}
    fp64_runs = {}  # config -> (inputs, FP64-form run) kept for a later F8 leg of the same config
    for name in also:
        if name in ("C4", "C4F8"):
            try:
                out["also"][name] = bench_c4(args.mixed if name == "C4" else "bf16f8", log_=log,
                                             graph=not args.c4_eager)
            except Exception as e:  # report, never hide; the headline line still prints
                out["also"][name] = {"error": repr(e)}
            continue
        if name == "C5":
            out["also"]["C5"] = bench_c5(log_=log)
            continue
        if name == "TW":
            out["also"]["TW"] = bench_tetwt(log_=log)
            continue
        if name == "IPW":
            out["also"]["IPW"] = bench_ipw(args.mixed, log_=log)
            continue
        if name == "GAM":
            out["also"]["GAM"] = bench_gamma(args.mixed, log_=log)
            continue
        if name == "ROT":
            out["also"]["ROT"] = bench_rotwave(log_=log)
            continue
        if name.endswith("F8") and (name[:-2] or args.config) in synth.CONFIGS:
            # the same q-point in the bf16f8 form (Sigma_c with scaled-fp8 cross terms)
            base = name[:-2] or args.config
            if base == args.config:
                q2, a64, st2 = qp, f64, args.steps
            elif base in fp64_runs:
                (q2, a64), st2 = fp64_runs.pop(base), 2
            else:
                log(f"bench: generating {base} inputs")
                q2 = synth.make_qpoint(synth.CONFIGS[base])
                a64, st2 = run_form(q2, "f64", 2, 1), 2
            try:
                a8 = run_form(q2, "bf16f8", st2, 1)
            except Exception as e:  # report, never hide; the headline line still prints
                out["also"][name] = {"error": repr(e)}
                continue
            f2 = flops_per_q(synth.CONFIGS[base], nwhis_pairs_bins=a8["npb"])
            out["also"][name] = {
                "workload": f"{base} q-point, precision bf16f8: bf16x3 except Sigma_c (bf16 hi*hi + scaled fp8 "
                            f"E4M3 cross terms)",
                "steps": st2, "s_per_q": round(a8["s_per_q"], 6),
                "speedup_vs_fp64_form": round(a64["s_per_q"] / a8["s_per_q"], 3),
                "qp_err_meV_vs_fp64_form": round(float(np.abs(a8["eqp"] - a64["eqp"]).max() * 1e3), 5),
                "tflops": round(sum(f2.values()) / a8["s_per_q"] / 1e12, 3),
                "kernel_ms_per_step": {k: round(v[0] / st2, 3) for k, v in a8["ktimes"].items()},
                "roofline": roofline(f2, a8, "bf16f8", st2, base),
                "tuning": a8["tuning"],
            }
            continue
        if name.endswith("R") and name[:-1] in synth.CONFIGS:
            out["also"][name] = bench_real_axis(args.mixed, name[:-1], log_=log)
            continue
        c = synth.CONFIGS[name]
        log(f"bench: generating {c.name} inputs")
        q2 = synth.make_qpoint(c)
        a64 = run_form(q2, "f64", 2, 1)
        amx = run_form(q2, args.mixed, 2, 1)
        f2 = flops_per_q(c, nwhis_pairs_bins=amx["npb"])
        out["also"][name] = {
            "workload": f"{c.name} q-point: nk={c.nk}, nb={c.nb}, occ={c.no}, unocc={c.nu}, nI={c.ni}, "
                        f"nf={c.nf}(+nu=0), ntq={c.nq}",
            "steps": 2, "warmup": 1,
            "s_per_q": round(amx["s_per_q"], 6), "fp64_form_s_per_q": round(a64["s_per_q"], 6),
            "speedup_vs_fp64_form": round(a64["s_per_q"] / amx["s_per_q"], 3),
            "qp_err_meV_vs_fp64_form": round(float(np.abs(amx["eqp"] - a64["eqp"]).max() * 1e3), 5),
            "tflops": round(sum(f2.values()) / amx["s_per_q"] / 1e12, 3),
            "kernel_ms_per_step": {k: round(v[0] / 2, 3) for k, v in amx["ktimes"].items()},
            "roofline": roofline(f2, amx, args.mixed, 2, name),
        }
        if name + "F8" in also:
            fp64_runs[name] = (q2, a64)
        del q2
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

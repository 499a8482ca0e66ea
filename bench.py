#!/usr/bin/env python3
"""Benchmark: LAPW H+O setup k-points/s on one MI355X.

One step = match + H/O setup for every k-point of the workload (default: workload 3, the
32-atom cell with 4 species, lmax 10 and 8 k-points of N_G+k about 5000, on which the
split-precision contraction is compared with FP64), with all inputs resident in HBM.  Prints one JSON line (the contract of the run driver):
metric/value/unit, the roofline of the dominant kernel (the muffin-tin MFMA contraction,
timed with HIP events on its own stream inside the library) and a CPU baseline (the FP64
oracle, timed on the host's cores on a bounded sample of the same k-points).
"""
import argparse
import json
import os
import pathlib
import sys
import time

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "H+O setup k-points/sec, 1 MI355X, split-BF16 vs FP64 at <=1e-8 Ha eig error"
PEAK_FP64 = 78.6e12          # MI355X dense FP64 MFMA, flop/s
PEAK_HBM = 8.0e12            # bytes/s

WORKLOAD_NAMES = {
    0: "cubic a=10.26, 2 atoms/1 species, lmax 8, 8 lo/atom, 1 k-point, |G+k|<3.5/R",
    1: "orthorhombic 8 atoms/2 species (R 2.0,1.8), lmax 10, apword 2, 4 lo/atom, 16 k-points, N_G+k in [1150,1250]",
    2: "16 atoms/3 species, lmax 10, 4 lo/atom, 64 k-points, N_G+k in [2400,2700]",
    3: "32 atoms/4 species, lmax 10, 4 lo/atom, 8 k-points, N_G+k ~5000",
    4: "16 atoms/3 species, lmax 10, 4 lo/atom, 8 k-points, N_G+k ~2500",
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def mt_flops(cr, n):
    """Algorithmic FLOPs of the MT contraction kernel per k: upper triangles of A^H.Y and A^H.A."""
    return 8.0 * cr.maxaa * cr.natm * n * n


def ystage_flops(cr, n):
    return 8.0 * cr.maxaa * cr.maxaa * cr.natm * n


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(w, budget_s=15.0):
    """The FP64 oracle (numpy/OpenBLAS on the host cores) on the first k-points of the workload,
    with the muffin-tin sums in the reference's own decomposition (one zgemm per atom,
    hamiltonsetup.f90:75-103, overlapsetup.f90:95-108) and match as the oracle restates it."""
    from oracle import lapw_ref as R
    try:
        from threadpoolctl import threadpool_info
        threads = max([d.get("num_threads", 1) for d in threadpool_info() if d.get("user_api") == "blas"] or [1])
    except Exception:
        threads = os.cpu_count() or 1
    t0 = time.perf_counter()
    done = 0
    for kp in w.kpoints:
        R.setup(w.crystal, kp, per_atom=True)
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return {"value": done / dt, "unit": "k-points/s", "cores": int(threads), "cpu_model": cpu_model(),
            "kind": "port",
            "sample": f"match + hamiltonsetup + overlapsetup of the first {done} k-point(s) of the same workload "
                      f"(numpy + host OpenBLAS, one zgemm per atom as the reference, {dt:.1f} s, no warm-up "
                      f"pass)"}


PEAK_BF16 = 2.5e15            # MI355X dense BF16 MFMA, flop/s


def bsplit_alg_flops(cr, ngps):
    """Algorithmic BF16 MFMA flops of the split contraction: for the upper triangles of H and O,
    6 slice pairs x 4 real products over the real K = natm * maxaa (tile and K padding excluded)."""
    return sum(6.0 * mt_flops(cr, n) for n in ngps)


PEAK_INT8 = 5.0e15            # MI355X dense INT8 MFMA, op/s


def lowest_eigs(H, O, nev=200):
    """The lowest nev eigenvalues of H c = e O c (LAPACK zhegvx on the host, range 'I', upper
    triangles, as solvewithlapack calls it)."""
    import scipy.linalg
    m = min(nev, H.shape[0])
    return scipy.linalg.eigh(H, O, lower=False, subset_by_index=[0, m - 1], driver="gvx", eigvals_only=True)


def crt_alg_ops(cr, ngps, nmod, share):
    """Algorithmic INT8 MFMA ops of the residue products: per modulus two real products (the
    two maps of Z_m[i]) over the upper triangle (N (N + 1) / 2 elements) and K = natm * maxaa,
    two ops per multiply-add; share = 0.5 when O comes from the addition theorem (H only)."""
    K = cr.natm * cr.maxaa
    return sum(2.0 * 2 * nmod * K * n * (n + 1) / 2 for n in ngps) * (2 * share)


def run_split(args, ctx, batch, cr, w, fp64_value, e0, log, olpflag=0, kind="crt"):
    """The split-precision leg: the muffin-tin contraction from exact integer residue products
    (kind 'crt', EXC_BATCH_SPLIT_CRT: the library's form and moduli count unless --moduli
    pins the count) or from three BF16 slices (kind
    'bf16', EXC_BATCH_SPLIT_BF16); everything else as the FP64 leg.  Eigenvalue error: the
    lowest 200 of k-point 0 against the FP64 GPU path's (e0)."""
    import torch
    from exciting_amd import lapw
    nk = batch.nk
    flags = lapw.EXC_BATCH_MATCH | olpflag | (lapw.EXC_BATCH_SPLIT_CRT if kind == "crt" else lapw.EXC_BATCH_SPLIT_BF16)
    tuned = None
    for kv in args.set:
        st, v = kv.split("=")
        ctx.set_variant(st, int(v))
        tuned = dict(tuned or {}, **{st: int(v)})
    if kind == "crt":
        ctx.set_moduli(args.moduli or 0)
    st = "crtgemm" if kind == "crt" else "bsplit"
    if args.tune:
        from exciting_amd import autotune
        tuned = autotune.tune(ctx, [st], lambda: ctx.setup_batched_dev(batch, flags), reps=2, log=log)
        if args.tune_out:
            autotune.save(args.tune_out, autotune.key(cr, [kp.ngp for kp in w.kpoints]), tuned)
    elif not args.set and st in (args.table or {}):
        tuned = {st: args.table[st]}           # pinned from the tuned table by main()
    ctx.setup_batched_dev(batch, flags)          # eager pass: stage breakdown
    torch.cuda.synchronize()
    ctx.read_timing()
    ctx.set_timing(True)
    ctx.setup_batched_dev(batch, flags)
    ctx.set_timing(False)
    eager = ctx.read_timing()
    for k, (ms, cnt) in eager.items():
        if cnt:
            log(f"  {kind} split stage {k:8s}: {ms:8.3f} ms")
    gflags = flags | lapw.EXC_BATCH_GRAPH
    for _ in range(args.warmup):
        ctx.setup_batched_dev(batch, gflags)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.setup_batched_dev(batch, gflags)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    value = nk * args.steps / dt
    Hs = batch.H[0].cpu().numpy().T.copy()
    Os = batch.O[0].cpu().numpy().T.copy()
    e1 = lowest_eigs(Hs, Os)
    err = float(np.abs(e1 - e0).max())
    share = 0.5 if olpflag else 1.0           # H only when O comes from the addition theorem
    ngps = [kp.ngp for kp in w.kpoints]
    fl64 = share * sum(mt_flops(cr, n) for n in ngps)
    form = nmod = sigma = None
    if kind == "crt":
        form, nmod, sigma = ctx.crt_form()
        ms = eager["crtgemm"][0]
        ops = crt_alg_ops(cr, ngps, nmod, share)
        # variants 8-10 are the register-pipelined kernel; the library's default by basis size
        # (the tuned table pins one for every workload here)
        v = (tuned or {}).get("crtgemm", 8 if max(ngps) >= 2048 else 4)
        gname = "crt_gemm_pipe_kernel" if v in (8, 9, 10) else "crt_gemm_kernel"
        roof = {"bound": "mfma", "kernel": f"{gname} (residue products, {nmod} moduli, int8 MFMA)",
                "achieved": ops / (ms * 1e-3) / 1e12, "peak": PEAK_INT8 / 1e12, "unit": "TOP/s",
                "frac": ops / (ms * 1e-3) / PEAK_INT8, "fp64_equivalent_TFLOPs": fl64 / (ms * 1e-3) / 1e12,
                "per_launch_ops": ops, "traffic": None}
        # HBM traffic of the residue GEMM from the committed rocprofv3 summary of the default command
        prof = HERE / "profiles" / f"setup_kernels_workload{args.config}.json"
        if args.nk is None and not args.tune and not args.set and args.moduli is None and olpflag and prof.exists():
            rec = json.loads(prof.read_text()).get(gname)
            if rec:
                roof["traffic"] = rec["fetch_bytes"] + rec["write_bytes"]
                roof["traffic_source"] = str(prof.relative_to(HERE))
    else:
        ms = eager["bsplit"][0]
        ex = share * bsplit_alg_flops(cr, ngps)
        roof = {"bound": "mfma", "kernel": "bsplit2_kernel (grouped split-BF16 MT)",
                "achieved": ex / (ms * 1e-3) / 1e12, "peak": PEAK_BF16 / 1e12, "unit": "TFLOP/s",
                "frac": ex / (ms * 1e-3) / PEAK_BF16, "fp64_equivalent_TFLOPs": fl64 / (ms * 1e-3) / 1e12}
    log(f"  {kind} split: {value:.1f} k-points/s ({value / fp64_value:.2f}x FP64), lowest-200 eig err {err:.2e} Ha")
    return {"value": value, "unit": "k-points/s", "speedup_vs_fp64": value / fp64_value,
            "ms_per_step": 1000.0 * dt / args.steps, "eig_err_lowest200_Ha_k0": err, "lowest_eig_Ha_k0": float(e0[0]),
            "meets_1e-8_Ha": err <= 1e-8, "moduli": nmod, "form": {0: "A,Yt", 1: "gram"}.get(form),
            "gram_sigma": sigma if form == 1 else None, "roofline": roof,
            "stages_ms": {k: v for k, (v, c) in eager.items() if c}, "variants": tuned}


def run_eigen(args, ctx, batch, w, olpflag, nev, log):
    """Setup + generalized eigensolve of every k on the device (the seceqnfv sequence), H and O
    consumed in place; the CPU reference is LAPACK zhegvx (scipy) on the same H, O."""
    import scipy.linalg
    import torch
    from exciting_amd import lapw
    flags = lapw.EXC_BATCH_MATCH | olpflag | lapw.EXC_BATCH_GRAPH
    evs = [torch.zeros(nev, dtype=torch.float64, device="cuda") for _ in range(batch.nk)]
    vecs = [torch.zeros((nev, n), dtype=torch.complex128, device="cuda") for n in batch.nmat]

    def step(nk=batch.nk):
        ctx.setup_batched_dev(batch, flags)
        ctx.eigen_batched_dev(batch.nmat[:nk], batch.H[:nk], batch.O[:nk], nev, evs[:nk], vecs[:nk])

    tuned = None
    if args.tune:
        from exciting_amd import autotune
        tuned = autotune.tune(ctx, ["eigen"], lambda: step(min(2, batch.nk)), reps=1, log=log)
    step()
    torch.cuda.synchronize()
    ctx.read_timing()
    ctx.set_timing(True)
    # the solve grows as nmat^3: at nmat ~5000 one step of 8 k-points takes seconds
    steps = args.steps if max(batch.nmat) <= 3000 else 1
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ctx.set_timing(False)
    tim = ctx.read_timing()
    ms_eig, cnt = tim["eigen"]
    value = batch.nk * steps / dt
    # host LAPACK on the first k (bounded: one solve, and only below nmat 3000)
    cpu_ms = None
    if batch.nmat[0] <= 3000:
        ctx.setup_batched_dev(batch, flags)
        torch.cuda.synchronize()
        H = batch.H[0].cpu().numpy().T.copy()
        O = batch.O[0].cpu().numpy().T.copy()
        t1 = time.perf_counter()
        scipy.linalg.eigh(H, O, lower=False, subset_by_index=[0, nev - 1], driver="gvx")
        cpu_ms = 1000.0 * (time.perf_counter() - t1)
    ms_eig /= batch.nk     # one timed launch covers all k-points
    log(f"  setup + eigensolve (lowest {nev}): {value:.1f} k-points/s, eigensolve {ms_eig / max(cnt, 1):.2f} ms/k"
        + (f" (host LAPACK zhegvx {cpu_ms:.0f} ms/k)" if cpu_ms else ""))
    dmat = run_dmat(args, ctx, batch, w, nev, vecs, log)
    return {"value": value, "unit": "k-points/s", "nev": nev, "ms_per_step": 1000.0 * dt / steps, "steps": steps,
            "eigensolve_ms_per_k": ms_eig / max(cnt, 1),
            "solver": "rocSOLVER: k-points padded to one order, strided-batched Cholesky, inverse, reduction and "
                      "tridiagonalization; per-k divide and conquer on 4 streams; back-transform of "
                      "all k-points by 64-reflector blocks (strided-batched GEMMs)",
            "variants": tuned, "cpu_lapack_zhegvx_ms_per_k": cpu_ms, "density_matrix": dmat}


def run_hmlint(args, w, log):
    """The per-SCF producers on a context of their own: the muffin-tin integrals
    (exc_ctx_hmlint: hmlint + olprad) on seeded radial data of the workload's atoms (meshes of
    500 / 460 points, lmaxvr 6), the density from a seeded Hermitian density matrix on the same
    radial data (exc_genrhomt_dev), and Theta(G), Theta(r), V(G) (exc_ctx_gencfun + genveffig) on
    the workload's G box with a seeded V(r); with the CPU baseline on, the numpy oracle's time
    for the same calls beside them."""
    import torch
    from exciting_amd import lapw, synth
    cr = w.crystal
    rd = synth.make_radial(cr, nr=(500, 460), lmaxvr=6)
    ctx = lapw.Context(cr)
    ctx.hmlint(rd, out=False)
    torch.cuda.synchronize()
    ctx.read_timing()
    ctx.set_timing(True)
    reps = max(args.steps, 3)
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.hmlint(rd, out=False)
    dt = (time.perf_counter() - t0) / reps
    ctx.set_timing(False)
    ms, cnt = ctx.read_timing()["hmlint"]
    res = {"ms_per_call": 1000.0 * dt, "kernel_ms": ms / max(cnt, 1), "natm": cr.natm, "lmaxvr": rd.lmaxvr,
           "nrmt": [int(x) for x in rd.nrmt], "cpu_oracle_ms": None}
    if not args.no_cpu_baseline:
        from oracle import hmlint_ref as M
        t1 = time.perf_counter()
        M.hmlint(cr, rd)
        M.olprad(cr, rd)
        res["cpu_oracle_ms"] = 1000.0 * (time.perf_counter() - t1)
    log(f"  muffin-tin integrals (hmlint + olprad, {cr.natm} atoms): {res['ms_per_call']:.2f} ms per call, "
        f"kernels {res['kernel_ms']:.3f} ms" + (f", numpy oracle {res['cpu_oracle_ms']:.0f} ms"
                                                 if res["cpu_oracle_ms"] else ""))
    # the muffin-tin density from a density matrix on the same radial data (genrhomt)
    ws = ctx.dmat_wfsize()
    g = torch.Generator(device="cuda").manual_seed(7)
    x = torch.randn((cr.natm, ws, ws), dtype=torch.complex128, device="cuda", generator=g)
    dm = (x + x.transpose(1, 2).conj()).contiguous()
    lmv, nrm = (rd.lmaxvr + 1) ** 2, rd.spr.shape[0]
    rho = torch.zeros((cr.natm, nrm, lmv), dtype=torch.float64, device="cuda")
    ctx._call("exc_genrhomt_dev", ctx.h, dm.data_ptr(), rho.data_ptr())
    torch.cuda.synchronize()
    ctx.read_timing()
    ctx.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx._call("exc_genrhomt_dev", ctx.h, dm.data_ptr(), rho.data_ptr())
    torch.cuda.synchronize()
    dt3 = (time.perf_counter() - t0) / reps
    ctx.set_timing(False)
    ms3, _ = ctx.read_timing()["hmlint"]
    res["density_from_dm"] = {"ms_per_call": 1000.0 * dt3, "kernel_ms": ms3 / reps, "wfsize": ws}
    log(f"  muffin-tin density from the density matrix (genrhomt): {1000.0 * dt3:.3f} ms per call, "
        f"kernels {ms3 / reps:.3f} ms")
    # Theta(G), Theta(r) and V(G) of the G box (gencfun + genveffig, hipFFT)
    rng = np.random.default_rng(20261202)
    v = rng.standard_normal(cr.ngrtot)
    ctx.gencfun()
    ctx.genveffig(v)
    torch.cuda.synchronize()
    ctx.read_timing()
    ctx.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.gencfun()
        ctx.genveffig(v)
    dt2 = (time.perf_counter() - t0) / reps
    ctx.set_timing(False)
    ms2, cnt2 = ctx.read_timing()["scf"]
    grid = [int(cr.intgv[i, 1] - cr.intgv[i, 0] + 1) for i in range(3)]
    res["interstitial_tables"] = {"ms_per_call": 1000.0 * dt2, "kernel_ms": ms2 / reps,
                                  "fft_grid": grid, "ngvec": int(cr.ngvec), "cpu_oracle_ms": None}
    if not args.no_cpu_baseline:
        from oracle import lapw_ref as R
        t1 = time.perf_counter()
        cf, cir = R.gencfun(cr)
        R.genveffig(cr, v, cir)
        res["interstitial_tables"]["cpu_oracle_ms"] = 1000.0 * (time.perf_counter() - t1)
    it = res["interstitial_tables"]
    log(f"  Theta(G), Theta(r), V(G) (gencfun + genveffig, grid {grid}): {it['ms_per_call']:.2f} ms per call, "
        f"kernels + FFTs {it['kernel_ms']:.3f} ms" + (f", numpy oracle {it['cpu_oracle_ms']:.0f} ms"
                                                     if it["cpu_oracle_ms"] else ""))
    res["direct_density"] = run_rhovalk(args, w, ctx, reps, log)
    return res


def run_rhovalk(args, w, ctx, reps, log):
    """The direct-route muffin-tin density of one k-point (exc_rhovalk_dev) at the reference's
    default lmaxvr 8, lradstep 1: radial data re-made at lmaxvr 8 (exc_ctx_hmlint), apwalm of
    the first k-point, 200 seeded coefficient vectors, occupations in [0, 2].  Flops counted
    are those of the factored sum the library runs (A^T C for l <= lmaxvr, the transformed
    coefficients P, the per-point matrices R, their contraction with the radial functions
    and the forward transform); the reference's direct sum, 8 lmmaxvr^2 per (atom, mesh
    point, state), is reported beside it.  The CPU baseline is the numpy oracle on the first
    20 states, scaled by 10."""
    import torch
    from exciting_amd import synth
    from oracle import rhovalk_ref as V
    cr, kp = w.crystal, w.kpoints[0]
    lmaxvr, nst = 8, 200
    rd = synth.make_radial(cr, nr=(500, 460), lmaxvr=lmaxvr)
    ctx.hmlint(rd, out=False)
    apw = ctx.match(kp)
    rng = np.random.default_rng(20261203)
    nmat = kp.ngp + cr.nlotot
    C = np.asfortranarray(rng.standard_normal((nmat, nst)) + 1j * rng.standard_normal((nmat, nst)))
    occ = rng.uniform(0.0, 2.0, nst)
    zb, _, rf = V.genshtmat(lmaxvr)
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x.T)).cuda()   # Fortran order on the device
    a_d, c_d, zb_d, rf_d, occ_d = t(np.asfortranarray(apw)), t(C), t(zb), t(rf), torch.from_numpy(occ).cuda()
    nrm, lmv = rd.spr.shape[0], (lmaxvr + 1) ** 2
    rho = torch.zeros((cr.natm, nrm, lmv), dtype=torch.float64, device="cuda")

    def call():
        ctx.rhovalk_dev(1, zb_d, rf_d, kp.ngp, apw.shape[0], a_d, nst, nmat, c_d, occ_d, 0.25, rho)   # evecsv: the identity

    call()
    torch.cuda.synchronize()
    ctx.read_timing()
    ctx.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(reps):
        call()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    ctx.set_timing(False)
    ms, _ = ctx.read_timing()["scf"]
    npts = sum(int(rd.nrmt[cr.spidx[a]]) for a in range(cr.natm))
    nf = cr.apwordmax * (lmaxvr + 1) + cr.nlomax
    npair = nf * (nf + 1) // 2
    lo_m = max(sum(2 * int(cr.lorbl[i, s]) + 1 for i in range(cr.nlorb[s]) if cr.lorbl[i, s] <= lmaxvr)
               for s in range(cr.nspecies))
    flops = (8.0 * cr.natm * cr.apwordmax * lmv * nst * kp.ngp             # W = A^T C, l <= lmaxvr
             + 8.0 * cr.natm * lmv * (cr.apwordmax * lmv + lo_m) * nst      # P
             + 4.0 * cr.natm * npair * lmv * nst                            # R
             + npts * (3.0 * npair * lmv + 2.0 * lmv * lmv))                # contraction, forward transform
    direct = 8.0 * lmv * lmv * npts * nst + 8.0 * cr.natm * cr.apwordmax * lmv * nst * kp.ngp
    res = {"ms_per_k": 1000.0 * dt, "kernel_ms": ms / reps, "states": nst, "lmaxvr": lmaxvr,
           "tflops": flops / (ms / reps * 1e-3) / 1e12, "flops_per_k": flops, "direct_sum_flops_per_k": direct,
           "cpu_oracle_ms_per_k": None}
    if not args.no_cpu_baseline:
        t1 = time.perf_counter()
        V.rhovalk(cr, rd, apw, kp.ngp, C[:, :20], occ[:20], 0.25, sht=(zb, None, rf))
        res["cpu_oracle_ms_per_k"] = 10.0 * 1000.0 * (time.perf_counter() - t1)
    log(f"  direct muffin-tin density (rhovalk, {nst} states, lmaxvr {lmaxvr}): {res['ms_per_k']:.3f} ms per k-point, "
        f"kernels {res['kernel_ms']:.3f} ms, {res['tflops']:.1f} TFLOP/s"
        + (f", numpy oracle {res['cpu_oracle_ms_per_k']:.0f} ms per k-point (20 states x 10)"
           if res["cpu_oracle_ms_per_k"] else ""))
    return res


def run_dmat(args, ctx, batch, w, nst, vecs, log):
    """Muffin-tin density matrix of all k-points (gendmatmt) from the eigenvectors of the
    eigensolve, seeded occupations in [0, 2] and equal k-point weights."""
    import numpy as np
    import torch
    cr = w.crystal
    rng = np.random.default_rng(20261202)
    occs = [torch.from_numpy(rng.uniform(0.0, 2.0, nst)).cuda() for _ in range(batch.nk)]
    wk = [1.0 / batch.nk] * batch.nk
    ws = ctx.dmat_wfsize()
    dm = torch.zeros((cr.natm, ws, ws), dtype=torch.complex128, device="cuda")
    ctx.gendmatmt_batched_dev(batch, nst, vecs, occs, wk, dm)
    torch.cuda.synchronize()
    ctx.read_timing()
    ctx.set_timing(True)
    for _ in range(args.steps):
        ctx.gendmatmt_batched_dev(batch, nst, vecs, occs, wk, dm)
    torch.cuda.synchronize()
    ctx.set_timing(False)
    ms, cnt = ctx.read_timing()["dmat"]
    ms /= max(cnt, 1)
    naa = cr.apwordmax * cr.lmmaxapw
    # A_a^T C per k and atom, then the weighted outer products (full wfsize^2, as the reference)
    flops = sum(8.0 * cr.natm * naa * nst * kp.ngp for kp in w.kpoints) + 8.0 * cr.natm * ws * ws * nst * batch.nk
    tf = flops / (ms * 1e-3) / 1e12
    log(f"  density matrix (gendmatmt, {nst} states): {ms / batch.nk:.3f} ms/k, {tf:.1f} TFLOP/s "
        f"({tf * 1e12 / PEAK_FP64:.2f} of the FP64 MFMA peak)")
    # the interstitial density of the same states (genrhoir: batched backward FFTs on the G box)
    rho = torch.zeros(cr.ngrtot, dtype=torch.float64, device="cuda")
    ctx.genrhoir_batched_dev(batch, nst, vecs, occs, wk, rho)
    torch.cuda.synchronize()
    ctx.read_timing()
    ctx.set_timing(True)
    for _ in range(args.steps):
        ctx.genrhoir_batched_dev(batch, nst, vecs, occs, wk, rho)
    torch.cuda.synchronize()
    ctx.set_timing(False)
    ms_ir, cnt_ir = ctx.read_timing()["scf"]
    ms_ir /= max(cnt_ir, 1)
    grid = [int(cr.intgv[i, 1] - cr.intgv[i, 0] + 1) for i in range(3)]
    log(f"  interstitial density (genrhoir, {nst} states, grid {grid}): {ms_ir / batch.nk:.3f} ms/k")
    vjp = run_vjp(args, ctx, batch, cr, nst, vecs, occs, wk, log)
    return {"ms_per_k": ms / batch.nk, "tflops": tf, "frac_fp64_peak": tf * 1e12 / PEAK_FP64, "flops_per_launch": flops,
            "k_points": batch.nk, "states": nst, "wfsize": ws,
            "interstitial_density_ms_per_k": ms_ir / batch.nk, "fft_grid": grid, "ibs_force_adjoint": vjp}


def run_vjp(args, ctx, batch, cr, nst, vecs, occs, wk, log):
    """The adjoint of the setup (exc_ho_vjp_dev: forcek's IBS term, dE/dtau and dE/dV(G)) on
    the eigenvectors of the eigensolve: wh = w occ, wo = -w occ e with seeded e (the time does
    not depend on the values), every k-point of the batch per step."""
    import torch
    rng = np.random.default_rng(20261204)
    ev = [np.sort(rng.uniform(-2.0, 2.0, nst)) for _ in range(batch.nk)]
    wh = [(wk[i] * occs[i]).contiguous() for i in range(batch.nk)]
    wo = [(-wk[i] * occs[i] * torch.from_numpy(ev[i]).cuda()).contiguous() for i in range(batch.nk)]
    dt = torch.zeros((cr.natm, 3), dtype=torch.float64, device="cuda")
    dv = torch.zeros(cr.ngvec, dtype=torch.complex128, device="cuda")

    def step():
        for i, d in enumerate(batch.dks):
            ctx.vjp_dev(d.ngp, d.ngkmax, batch.apw[i], d.igpig, d.vgpc, d.sfacgp, vecs[i], batch.nmat[i], nst,
                        wh[i], wo[i], dt, dv)

    try:
        step()
        torch.cuda.synchronize()
    except Exception as e:   # reported, not fatal: the other legs' figures stand
        log(f"  IBS force adjoint: not measured ({e})")
        return None
    ctx.read_timing()
    ctx.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt_s = (time.perf_counter() - t0) / args.steps
    ctx.set_timing(False)
    ms, cnt = ctx.read_timing()["vjp"]
    res = {"ms_per_k": 1000.0 * dt_s / batch.nk, "kernel_ms_per_k": ms / max(cnt, 1), "states": nst}
    log(f"  IBS force adjoint (exc_ho_vjp_dev, {nst} states): {res['ms_per_k']:.3f} ms/k, "
        f"kernels {res['kernel_ms_per_k']:.3f} ms/k")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=3)
    ap.add_argument("--nk", type=int, default=None, help="limit the number of k-points")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tune", action="store_true", help="autotune tile variants before timing")
    ap.add_argument("--tune-out", default=None, metavar="JSON",
                    help="with --tune: merge the chosen variants into this table under the batch's shape key")
    ap.add_argument("--no-table", action="store_true",
                    help="ignore the committed tuned-variant table (exciting_amd/tuned.json); built-in defaults")
    ap.add_argument("--set", action="append", default=[], metavar="STAGE=VARIANT",
                    help="pin a stage's tile variant (e.g. bmt=1, ystage=2); repeatable")
    ap.add_argument("--no-split", action="store_true", help="skip the split-precision measurements")
    ap.add_argument("--no-bf16", action="store_true", help="skip the three-slice BF16 split measurement")
    ap.add_argument("--moduli", type=int, default=None,
                    help="moduli of the modular split (8..12; default: the library's, 10 with the Gram form, else 11)")
    ap.add_argument("--no-hmlint", action="store_true", help="skip the muffin-tin integrals (hmlint) measurement")
    ap.add_argument("--eigen", type=int, default=200, metavar="NEV",
                    help="grouped path: also time setup + device eigensolve of the lowest NEV states "
                         "(0: skip)")
    ap.add_argument("--no-olp", action="store_true",
                    help="grouped path: overlap muffin-tin part by the MFMA contraction instead of the "
                         "addition theorem (EXC_BATCH_OLP_LEGENDRE)")
    ap.add_argument("--path", choices=["batched", "perk", "host"], default="batched",
                    help="batched: all k-points in one grouped launch per stage, replayed as a hipGraph; "
                         "perk: one launch sequence per k-point on device arrays; host: the host-array "
                         "entry points the Fortran driver calls (PCIe copies of apwalm, H and O included)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 1 GiB device-to-device copy (calibrates rocprof FETCH/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); the path shards over "
              f"k-points across GPUs by design, but no multi-GPU run is part of it", file=sys.stderr)
        sys.exit(2)

    import torch
    from exciting_amd import lapw, synth

    if args.copy_probe:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")     # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("copy probe: 5 copies of 1073741824 bytes")
        return

    w = synth.make_config(args.config, nk=args.nk)
    cr = w.crystal
    nk = len(w.kpoints)
    log(f"workload {args.config}: natm={cr.natm} lmax={cr.lmaxapw} nk={nk} N={[k.ngp for k in w.kpoints]}")
    stream = torch.cuda.Stream()          # default priority; a non-null stream can be graph-captured
    torch.cuda.set_stream(stream)
    ctx = lapw.Context(cr, stream=stream)
    batch = lapw.DeviceBatch(cr, w.kpoints)
    dks, apw, nmats, Hs, Os = batch.dks, batch.apw, batch.nmat, batch.H, batch.O

    def step_perk():
        for i, d in enumerate(dks):
            ctx.match_dev(d.ngp, d.ngkmax, d.gpc, d.tpgpc, d.sfacgp, apw[i])
            ctx.setup_dev(d.ngp, d.ngkmax, apw[i], d.igpig, d.vgpc, nmats[i], Hs[i], Os[i])

    olpflag = lapw.EXC_BATCH_OLP_LEGENDRE if (args.path == "batched" and not args.no_olp) else 0

    def step_batched(graph=True):
        ctx.setup_batched_dev(batch, lapw.EXC_BATCH_MATCH | olpflag | (lapw.EXC_BATCH_GRAPH if graph else 0))

    if args.path == "host":
        # preallocated host arrays, as a Fortran caller holds them (pageable memory)
        h_apw = [np.zeros((kp.igpig.shape[0], cr.apwordmax, cr.lmmaxapw, cr.natm), np.complex128, order="F")
                 for kp in w.kpoints]
        h_ho = [(np.zeros((kp.ngp + cr.nlotot,) * 2, np.complex128, order="F"),
                 np.zeros((kp.ngp + cr.nlotot,) * 2, np.complex128, order="F")) for kp in w.kpoints]

    def step_host():
        # as the Fortran driver runs them: match fills the caller's apwalm, the setup reads the
        # library's device copy of it (the array is unchanged since), H and O come back
        for i, kp in enumerate(w.kpoints):
            ctx.match(kp, out=h_apw[i])
            ctx.hamiltonandoverlapsetup(kp, h_apw[i], out=h_ho[i], resident=True)

    tuned = None
    args.table = None
    if args.path == "batched" and not args.tune and not args.set and not args.no_table:
        # the committed table's variants for this batch shape (built-in defaults where it has none)
        from exciting_amd import autotune
        args.table = autotune.apply(ctx, cr, [kp.ngp for kp in w.kpoints]) or None
        if args.table:
            log(f"tuned table {autotune.key(cr, [kp.ngp for kp in w.kpoints])}: {args.table}")
            tuned = {k: v for k, v in args.table.items() if k in ("bmt", "ystage", "olp", "graph")} or None
    for kv in args.set:
        st, v = kv.split("=")
        ctx.set_variant(st, int(v))
        tuned = dict(tuned or {}, **{st: int(v)})
    if args.tune:
        from exciting_amd import autotune
        if args.path in ("perk", "host"):
            tuned = autotune.tune(ctx, ["mt", "ystage"], step_perk, reps=2, log=log)
        else:
            tuned = autotune.tune(ctx, ["bmt", "ystage"] + (["olp"] if olpflag else []), lambda: step_batched(False),
                                  reps=2, log=log)
            tuned.update(autotune.tune(ctx, ["graph"], lambda: step_batched(True), reps=3, log=log))
            if args.tune_out:
                autotune.save(args.tune_out, autotune.key(cr, [kp.ngp for kp in w.kpoints]), tuned)
    if args.path == "batched":
        # stage breakdown of one eager (uncaptured) grouped pass, then time the graph replays
        step_batched(False)
        torch.cuda.synchronize()
        ctx.read_timing()
        ctx.set_timing(True)
        step_batched(False)
        ctx.set_timing(False)
        eager = ctx.read_timing()
        for k, (ms, cnt) in eager.items():
            if cnt:
                log(f"  grouped stage {k:7s}: {ms:8.3f} ms")
        step = step_batched
    else:
        eager = None
        step = step_perk if args.path == "perk" else step_host
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ctx.read_timing()
    ctx.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ctx.set_timing(False)
    tim = ctx.read_timing()
    elapsed = t1 - t0
    value = nk * args.steps / elapsed
    ms_step = 1000.0 * elapsed / args.steps
    for k, (ms, cnt) in tim.items():
        if cnt:
            log(f"  stage {k:7s}: {ms / cnt:8.3f} ms/launch over {cnt} launches")

    # MFMA flops of the dominant kernel: the H and O contractions, or H alone when O's
    # muffin-tin part comes from the addition theorem
    share = 0.5 if olpflag else 1.0
    if args.path == "batched":
        mt_ms, mt_cnt = eager["bmt"]
        fl = share * sum(mt_flops(cr, kp.ngp) for kp in w.kpoints)
    else:
        mt_ms, mt_cnt = tim["mt"]
        fl = sum(mt_flops(cr, kp.ngp) for kp in w.kpoints) * args.steps
    achieved = fl / (mt_ms * 1e-3) if mt_ms > 0 else 0.0
    # the grouped MT variants 0, 2, 3 (0 is the default) form each complex product from three
    # real MFMA products (Gauss's form) instead of four: they execute 6 of the 8 algorithmic flops
    bmt_v = (tuned or {}).get("bmt", 0)
    gauss = args.path == "batched" and bmt_v in (0, 2, 3)
    roof = {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64, "traffic": None,
            "kernel": ("bmt_dma_grid_kernel (grouped, all k)" if args.path == "batched" else "mt_dma_kernel (per k)")
                      + (": A^H.Y + interstitial (H)" if olpflag else ": A^H.Y and A^H.A + interstitial")
                      + (", 3 real MFMA products per complex product" if gauss else ""),
            "per_launch_flops": fl / max(mt_cnt, 1),
            "executed_mfma_frac": (0.75 if gauss else 1.0) * achieved / PEAK_FP64}
    # HBM traffic of the dominant kernel: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this same
    # default command, scaled by a 1 GiB copy (committed summary); only for the profiled setup
    prof = HERE / "profiles" / f"setup_kernels_workload{args.config}.json"
    if (args.path == "batched" and args.nk is None and not args.tune and not args.set and olpflag
            and prof.exists()):
        rec = json.loads(prof.read_text()).get("bmt_dma_grid_kernel")
        if rec:
            roof["traffic"] = rec["fetch_bytes"] + rec["write_bytes"]
            roof["traffic_source"] = str(prof.relative_to(HERE))
            roof["algorithmic_bytes"] = sum(32.0 * kp.ngp * cr.natm * cr.maxaa + 16.0 * kp.ngp ** 2
                                            for kp in w.kpoints)
    stages = {k: {"ms_per_launch": (ms / cnt if cnt else None), "launches": cnt} for k, (ms, cnt) in tim.items() if cnt}
    if eager:
        stages["grouped_eager_pass"] = {k: ms for k, (ms, cnt) in eager.items() if cnt}

    split = None
    if args.path == "batched" and not args.no_split:
        # eigenvalues of the FP64 result of k-point 0, the split legs' yardstick
        step_batched(False)
        torch.cuda.synchronize()
        e0 = lowest_eigs(batch.H[0].cpu().numpy().T.copy(), batch.O[0].cpu().numpy().T.copy())
        split = {"crt": run_split(args, ctx, batch, cr, w, value, e0, log, olpflag, "crt")}
        if not args.no_bf16:
            split["bf16"] = run_split(args, ctx, batch, cr, w, value, e0, log, olpflag, "bf16")
    eig = None
    if args.path == "batched" and args.eigen > 0:
        eig = run_eigen(args, ctx, batch, w, olpflag, min(args.eigen, min(batch.nmat)), log)

    hml = None
    if args.path == "batched" and not args.no_hmlint:
        hml = run_hmlint(args, w, log)
    base = None
    if not args.no_cpu_baseline:
        base = cpu_baseline(w)
    out = {
        "metric": METRIC, "value": value, "unit": "k-points/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded generator, no checkpoints)",
        "config": {"workload": f"workload {args.config}: " + WORKLOAD_NAMES[args.config], "k_points": nk,
                   "path": ("FP64, all k-points grouped per stage, hipGraph replay"
                            + (", overlap MT part by the addition theorem" if olpflag else "")
                            if args.path == "batched" else "FP64, per-k launches"
                            + (", host arrays (PCIe copies included)" if args.path == "host" else "")),
                   "parallelism": "single GPU", "variants": tuned},
        "roofline": roof, "stages": stages, "split": split, "setup_and_eigensolve": eig,
        "muffin_tin_integrals": hml, "cpu_baseline": base,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

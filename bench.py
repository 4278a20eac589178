"""Throughput of the simulated-bifurcation step on one MI355X.

Workload (default): MaxCut dSB on a random 4-regular graph, N = 2^20 nodes,
Z = 4 N = 4,194,304 symmetric nonzeros, weights +-1 uniform from seed, 256
agents, dt = 0.5, max_steps = K (20,000 by default), windows every 100 steps.
A "step" is one fused SB step over all N x 256 node-agents.  The headline
form (--form, default q8s: X and Y stored as 8-bit fixed-point codes, 2
bytes per node-agent, Y rounded stochastically with a hashed dither, FP32
arithmetic, J sgn(X) exact from packed sign words) runs beside the plain
FP32 dSB kernel for the speed-up and the equal-steps best-cut comparison;
--also adds more forms (default: the 10- and 12-bit forms q10 and q12; the
8-bit form with nearest-even rounding q8, outside the 0.1% best-cut bar on
some graph families -- see --workload quality; and the packed-sign form,
bit-identical to plain FP32), and the line also carries bSB with BF16
gathers against plain FP32 bSB, a batched-serving summary (4096 small
instances) and the CPU baseline.

The timed region is one sis_solver_run (pre-step + steps + windows), with the
state resident in HBM; the update window's early exit is off inside the timed
region.  A run executes whole window periods (100 steps), so K is rounded up
to a multiple of 100: the line's "steps" is the executed count and
ms_per_step divides the timed region by it.  The step-kernel time comes from
HIP events the library records around the step launches on the solver's
stream.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--form packed|f32|bf16|f64]
                    [--baseline-form f32] [--also f1,f2] [--cpu-seconds S] [--no-cpu]
                    [--no-bf16] [--no-serving] [--workload ...]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sparse-ising-solver_amd", "python"))

import numpy as np  # noqa: E402

import sis  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "agent-steps/s + HBM GB/s % of peak; best cut vs plain FP32 SB at equal steps"


def step_bytes(form, N, Z, A):
    """Algorithmic HBM bytes of one step launch over all agents."""
    csr = 8 * (N + 1) + 4 * Z + 4 * Z  # row_ptr int64, col int32, val fp32
    state = 16 * N * A                 # X, Y read + write, FP32
    if form == "packed":
        return state + N * A // 8 + Z * A // 8 + csr
    if form == "q12":  # 12-bit X and Y codes: 3 bytes read + 3 written
        return 6 * N * A + N * A // 8 + Z * A // 8 + csr
    if form == "q10":  # 10-bit X and Y codes: 2.5 bytes read + 2.5 written
        return 5 * N * A + N * A // 8 + Z * A // 8 + csr
    if form in ("q8", "q8s"):  # 8-bit X and Y codes: 2 bytes read + 2 written
        return 4 * N * A + N * A // 8 + Z * A // 8 + csr
    if form == "bf16":
        return state + 2 * N * A + 2 * Z * A + csr
    if form == "f64":
        return 32 * N * A + 8 * Z * A + 8 * (N + 1) + 4 * Z + 8 * Z
    return state + 4 * Z * A + csr      # plain f32


def measured_traffic(form, tile, N, Z, A):
    """HBM bytes per launch of this kernel tile from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE runs (calibrated on k_prestep), or None."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_maxcut_4regular_N2p20_A256.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except OSError:
        return None
    if (t["nodes"], t["nnz"], t["agents"]) != (N, Z, A):
        return None
    k = t["kernels"].get(f"{form}:{tile}")
    return k["traffic_bytes"] if k else None


def time_form(J, form, A, dt, K, W, seed, dyn, autotune=True):
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=A, dt=dt, form=sis.FORMS[form], dynamics=dyn))
    tune = None
    if autotune:  # tile choice, outside the timed region; every tile is bit-identical
        # 50-step trials: the narrow forms' best two tiles are within a few
        # percent, which 10-step trials do not resolve reliably
        s.init_random(seed + 1)
        tile, ms = s.autotune(50)
        tune = dict(tile=tile, trial_ms_per_step=ms)
    if W > 0:  # untimed warm-up steps (same kernels, discarded state)
        s.init_random(seed + 1)
        s.advance(0, W, max(K, 1))
    s.init_random(seed)
    t0 = time.perf_counter()
    info = s.run(K, 0)
    wall = time.perf_counter() - t0
    cuts, best = s.cuts()
    check = energy_check(J, s, cuts, best)
    cont = continuous_energy(J, s) if dyn == sis.BSB else None
    s.close()
    return dict(info=info, wall=wall, best_cut=float(cuts.max()), best_agent=int(best),
                mean_cut=float(cuts.mean()), tune=tune, energy_check=check, cont_energy=cont)


def graph_run(J, form, A, dt, K, seed, dyn, tile, head):
    """The headline solve launched as one captured hipGraph per run
    (use_graph 2: every window period's step and window kernels), replayed.
    A first run captures and instantiates the graph (untimed); the second,
    from the same seed, is timed.  Its results must equal the direct-launch
    run's bit for bit."""
    p = dict(problem=sis.MAXCUT, agents=A, dt=dt, form=sis.FORMS[form], dynamics=dyn, use_graph=2)
    if tile is not None:
        p["tile"] = tile
    s = sis.Solver(J, p)
    s.init_random(seed)
    s.run(K, 0)
    s.init_random(seed)
    t0 = time.perf_counter()
    info = s.run(K, 0)
    wall = time.perf_counter() - t0
    cuts, best = s.cuts()
    s.close()
    ex = max(info["steps_executed"], 1)
    value = info["steps_executed"] * A / wall
    return dict(use_graph=2, launch="one hipGraph per run: every window period's step and "
                "window kernels, captured once and replayed", steps_executed=info["steps_executed"],
                value=value, ms_per_step=wall * 1e3 / ex, device_ms=info["device_ms"],
                speedup_vs_direct=value / head["value"], best_cut=float(cuts.max()),
                identical_to_direct=bool(float(cuts.max()) == head["best_cut"]
                                         and float(cuts.mean()) == head["mean_cut"]))


def continuous_energy(J, s, nagents=16):
    """bSB: energy (1/2) x^T J x of the continuous state X after the run, for
    the first agents, evaluated in FP64 on the host (scipy)."""
    import scipy.sparse as sp

    rp, col, val = J.arrays()
    M = sp.csr_matrix((val, col, rp), shape=(J.n, J.n))
    x = s.get_state()[0][:, :min(nagents, s.A)]
    return 0.5 * np.einsum("ia,ia->a", x, M @ x)


def energy_check(J, s, cuts, best, nagents=16):
    """Device FP64 cut and energy of the returned spins against an FP64 host
    evaluation (scipy) of the same spins, for the best agent and the first
    agents; the error is 0 for integer weights."""
    import scipy.sparse as sp

    rp, col, val = J.arrays()
    M = sp.csr_matrix((val, col, rp), shape=(J.n, J.n))
    idx = sorted(set([int(best)] + list(range(min(nagents, s.A)))))
    spins = s.spins(sis.SPINS_CURRENT)[:, idx].astype(np.float64)
    e_host = np.einsum("ia,ia->a", spins, M @ spins)      # s^T J s
    cut_host = (J.sum - 0.5 * e_host) / 2                  # -(1/2)(H - S), H = s^T J s / 2
    e_dev = s.energy(sis.SPINS_CURRENT)[idx]
    return dict(agents=len(idx), best_agent_included=True,
                max_abs_err_cut=float(np.abs(cut_host - cuts[idx]).max()),
                max_abs_err_energy=float(np.abs(e_host - e_dev).max()),
                host="scipy FP64 sparse product of the device's current spins")


def host_cpu():
    """lscpu model, sockets and physical cores of the host."""
    import subprocess

    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, check=True).stdout
        kv = dict(line.split(":", 1) for line in out.splitlines() if ":" in line)
        kv = {k.strip(): v.strip() for k, v in kv.items()}
        sockets = int(kv.get("Socket(s)", "1"))
        per = int(kv.get("Core(s) per socket", "0"))
        info = dict(model=kv.get("Model name", "?"), sockets=sockets,
                    physical_cores=sockets * per,
                    threads_per_core=int(kv.get("Thread(s) per core", "1")))
    except (OSError, ValueError, subprocess.CalledProcessError):
        pass
    return info


def native_oracle():
    """The CPU step built for this host: oracle/sis_oracle.c (rows in parallel,
    agents innermost so they vectorise) with -O3 -march=native -fopenmp,
    compiled into a temporary directory.  Falls back to the prebuilt
    liboracle.so (-O3 -mavx2 -mfma) when no compiler is available."""
    import subprocess
    import tempfile

    src = os.path.join(ROOT, "oracle", "sis_oracle.c")
    out = os.path.join(tempfile.mkdtemp(prefix="sis_cpu_"), "liboracle_native.so")
    cmd = ["gcc", "-O3", "-march=native", "-fPIC", "-std=c11", "-ffp-contract=off", "-fopenmp",
           "-shared", "-o", out, src, "-lm"]
    try:
        subprocess.run(cmd, check=True, capture_output=True)
        return out, " ".join(cmd[:8])
    except (OSError, subprocess.CalledProcessError):
        return os.path.join(ROOT, "oracle", "liboracle.so"), "prebuilt -O3 -mavx2 -mfma"


def cpu_baseline(J, A, dt, xi0, seconds, seed):
    """FP32 dSB step (the plain form's arithmetic) on the host cores: the
    oracle's step compiled natively, timed over a bounded sample of steps of
    the same graph and agents (steady_clock-equivalent perf_counter around the
    step loop only)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    path, build = native_oracle()
    orc.LIB_PATH = path
    orc._lib = None
    rp, col, val = J.arrays()
    pb = orc.Problem(rp, col, val, S=J.sum)
    X, Y = orc.init_state(seed, orc.MAXCUT, J.n, A)
    T = 20000
    threads = orc.max_threads()

    def timed(t0, seconds):
        c = time.perf_counter()
        orc.advance(pb, orc.MAXCUT, orc.DSB, X, Y, t0, 1, T, dt, xi0, 1.0, "f32")
        first = time.perf_counter() - c
        n = max(1, int(seconds / max(first, 1e-6)) - 1)
        c = time.perf_counter()
        orc.advance(pb, orc.MAXCUT, orc.DSB, X, Y, t0 + 1, n, T, dt, xi0, 1.0, "f32")
        return n, time.perf_counter() - c

    n, el = timed(0, seconds)
    orc.set_threads(1)
    c = time.perf_counter()
    orc.advance(pb, orc.MAXCUT, orc.DSB, X, Y, n + 1, 2, T, dt, xi0, 1.0, "f32")
    el1 = (time.perf_counter() - c) / 2
    orc.set_threads(threads)
    return dict(value=n * A / el, unit="agent-steps/s", cores=threads, kind="port",
                one_core=A / el1, host=host_cpu(), build=build,
                sample=f"FP32 dSB step of the same graph and agents (rows in parallel, agents "
                       f"innermost), {threads} OpenMP threads (the process's CPU share), "
                       f"{n} steps after 1 warm-up step in {el:.1f} s; one_core: 2 steps on "
                       f"1 thread, {2 * el1:.1f} s")


def run_solver(J, params, T, uw, seed, field=None, autotune=True, tsp=None):
    """Timed sis_solver_run (pre-step + T steps + windows) of one configuration.
    With params["structured"] the TSP solver is built from `tsp` directly (no
    matrix)."""
    if params.get("structured"):
        s = sis.Solver.from_tsp(tsp, params)
    else:
        s = sis.Solver(J, params, field=field)
    tune = None
    if autotune:
        s.init_random(seed + 1)
        tune = s.autotune(10)[0]
    s.init_random(seed)
    s.run(min(T, 200), uw)  # warm-up run (graph capture, caches); discarded
    s.init_random(seed)
    t0 = time.perf_counter()
    info = s.run(T, uw)
    wall = time.perf_counter() - t0
    out = dict(info=info, wall=wall, tile=tune, split=s.p.split,
               agent_steps_per_s=info["steps_executed"] * s.A / wall)
    if tsp is None:
        cuts, best = s.cuts()
        out.update(best=float(cuts.max()), mean=float(cuts.mean()))
    else:
        lens, best = s.tsp_lengths(tsp)
        ok = lens[lens >= 0]
        out.update(valid=int(len(ok)), best=float(ok.min()) if len(ok) else None,
                   mean=float(ok.mean()) if len(ok) else None)
        out.update(tsp_energies(s, lens))
    s.close()
    return out


def tsp_energies(s, lens):
    """Energies s^T J s - h.s of the spins a TSP run returns: the minimum over
    every agent's current spins (sgn X at the last window), and over the final
    spins of agents that reached a valid tour (agents without one keep the
    reset value there, so they are left out)."""
    Ec = s.energy(sis.SPINS_CURRENT)
    out = dict(energy_min_current=float(Ec.min()))
    if (lens >= 0).any():
        out["energy_min_final_valid"] = float(s.energy(sis.SPINS_FINAL)[lens >= 0].min())
    return out


def small_configs(a):
    """The remaining configurations, one JSON line each (not the headline)."""
    lines = []
    if a.workload == "er800":
        # ER G(800, 0.06), +1 weights, 128 agents, bSB, dt 0.1, 2000 steps, no
        # window exit; c0 = 0.5 / (sqrt(mean degree) * std(J)) as configured.
        J = sis.Csr.gen_er(800, 0.06, sis.W_ONE, a.seed)
        mean_deg = J.nnz / J.n
        sigma = 0.7 / (J.xi0() * np.sqrt(J.n))  # sample std of all N^2 entries
        c0 = 0.5 / (np.sqrt(mean_deg) * sigma)
        # the configured c0, then the reference's own xi0 = 0.7 / (sigma sqrt N)
        for g, xi0, kind in ((0, c0, "configured"), (1, c0, "configured"), (2, c0, "configured"),
                             (1, J.xi0(), "reference")):
            r = run_solver(J, dict(problem=sis.MAXCUT, agents=128, dt=0.1, dynamics=sis.BSB,
                                   form=sis.FORM_F32, xi0=xi0, use_graph=g), 2000, 0, a.seed)
            lines.append(dict(workload="ER N=800 p=0.06 bSB plain FP32, 128 agents, 2000 steps",
                              use_graph=g, nnz=J.nnz, xi0=xi0, xi0_kind=kind,
                              **{k: v for k, v in r.items() if k != "info"},
                              ms_per_step=r["wall"] * 1e3 / r["info"]["steps_executed"]))
    elif a.workload == "torus100":
        J = sis.Csr.gen_torus(100, 100, sis.W_PM1, a.seed)
        for form in ("f32", "packed", "q12", "q10", "q8", "q8s"):
            r = run_solver(J, dict(problem=sis.MAXCUT, agents=1024, dt=0.5, form=sis.FORMS[form],
                                   use_graph=1), a.steps if a.steps != 20000 else 10000, 1, a.seed)
            ms = r["info"]["step_kernel_ms"] / max(r["info"]["steps_executed"], 1)
            lines.append(dict(workload="torus 100x100 +-1 dSB, 1024 agents, 10000 steps, uw=1",
                              form=form, steps_executed=r["info"]["steps_executed"],
                              **{k: v for k, v in r.items() if k != "info"},
                              ms_per_step=r["wall"] * 1e3 / max(r["info"]["steps_executed"], 1)))
    elif a.workload == "tsp48":
        t = sis.Tsp.generate(48, 1000, 1, a.seed)
        J, h = t.qubo()
        # CSR forms with and without the row split, then the structured
        # product (no matrix, distance GEMM on the matrix cores)
        for form, split, st in (("f32", 0, 0), ("bf16", 0, 0), ("f32", -1, 0), ("bf16", -1, 0),
                                ("f32", 0, 1)):
            r = run_solver(J, dict(problem=sis.TSP, agents=1024, dt=0.1, form=sis.FORMS[form],
                                   dynamics=sis.BSB, ncities=48, use_graph=1, split=split,
                                   structured=st),
                           a.steps if a.steps != 20000 else 10000, 0, a.seed, field=h, tsp=t)
            lines.append(dict(workload="TSP n=48 (2304 spins) bSB, 1024 agents, 10000 steps",
                              form=form, split_requested=split, structured=st, nnz=J.nnz,
                              **{k: v for k, v in r.items() if k != "info"},
                              ms_per_step=r["wall"] * 1e3 / r["info"]["steps_executed"],
                              step_kernel_ms=r["info"]["step_kernel_ms"]))
    elif a.workload == "tsp_struct":
        # TSP bSB FP32 with the QUBO's CSR against the structured product (no
        # matrix): n = 48 (the 10,000-step configuration), 128 and 256 cities;
        # n = 512 structured only (its CSR would hold 537M entries).
        for n, A, T in ((48, 1024, 10000), (128, 256, 1000), (256, 256, 1000), (512, 256, 500)):
            t = sis.Tsp.generate(n, 1000, 1, a.seed)
            for structured in ((1, 0) if n <= 256 else (1,)):
                params = dict(problem=sis.TSP, agents=A, dt=0.1, form=sis.FORM_F32,
                              dynamics=sis.BSB, structured=structured)
                t0 = time.perf_counter()
                s = sis.Solver.from_tsp(t, params)
                build = time.perf_counter() - t0
                s.init_random(a.seed)
                s.run(min(T, 200), 0)
                s.init_random(a.seed)
                t0 = time.perf_counter()
                info = s.run(T, 0)
                wall = time.perf_counter() - t0
                lens, _ = s.tsp_lengths(t)
                ok = lens[lens >= 0]
                lines.append(dict(workload=f"tsp bSB f32, n={n} ({n * n} spins), {A} agents, {T} steps",
                                  structured=structured, create_s=build, wall=wall,
                                  ms_per_step=wall * 1e3 / info["steps_executed"],
                                  agent_steps_per_s=info["steps_executed"] * A / wall,
                                  valid=int(len(ok)), best=float(ok.min()) if len(ok) else None,
                                  **tsp_energies(s, lens)))
                s.close()
    elif a.workload == "tsp_valid":
        # Valid-tour rate of the one-hot TSP QUBO against the number of cities
        # and the coordinate span, with the reference's fixed coupling 0.002
        # (src/ising_tsp.cu:61) and its alternative 0.5 / (sigma sqrt N)
        # (src/ising_tsp.cu:62-83): bSB FP32, the reference's 10,000 agents,
        # 10,000 steps, dt 0.1, window every 1,000 steps; structured product
        # on resident workgroups (tile 3).
        T = a.steps if a.steps != 20000 else 10000
        A = a.agents if a.agents != 256 else 10000
        for n in (8, 10, 12, 14, 16, 20, 24, 32, 48):
            for span in (100, 1000):
                t = sis.Tsp.generate(n, span, 1, a.seed + n)
                for rule in ("fixed", "scaled"):
                    xi0 = 0.002 if rule == "fixed" else t.xi0()
                    s = sis.Solver.from_tsp(t, dict(agents=A, form=sis.FORM_F32, dynamics=sis.BSB,
                                                    dt=0.1, xi0=xi0, structured=1, tile=3,
                                                    use_graph=1))
                    s.init_random(a.seed)
                    t0 = time.perf_counter()
                    info = s.run(T, 0)
                    wall = time.perf_counter() - t0
                    lens, _ = s.tsp_lengths(t)
                    ok = lens[lens >= 0]
                    lines.append(dict(workload=f"tsp bSB f32 structured resident, {A} agents, "
                                               f"{T} steps, dt 0.1", n=n, span=span,
                                      xi0_rule=rule, xi0=xi0, valid=int(len(ok)),
                                      valid_fraction=len(ok) / A,
                                      best=float(ok.min()) if len(ok) else None,
                                      mean=float(ok.mean()) if len(ok) else None,
                                      random_tour=t.random_length(a.seed),
                                      steps_executed=info["steps_executed"], wall_s=wall,
                                      **tsp_energies(s, lens)))
                    s.close()
        # The same against c = xi0 n max(d), the scale of xi0 times the field
        # (h ~ (n - 2)(B + C) / 2 with B = C = max d): 1,000-span instances.
        for n in (16, 32, 48, 64):
            t = sis.Tsp.generate(n, 1000, 1, a.seed + n)
            for c in (10, 15, 20, 30, 45, 60, 90):
                xi0 = c / (n * t.maxd)
                s = sis.Solver.from_tsp(t, dict(agents=A, form=sis.FORM_F32, dynamics=sis.BSB,
                                                dt=0.1, xi0=xi0, structured=1, tile=3,
                                                use_graph=1))
                s.init_random(a.seed)
                info = s.run(T, 0)
                lens, _ = s.tsp_lengths(t)
                ok = lens[lens >= 0]
                lines.append(dict(workload=f"tsp bSB f32 structured resident, {A} agents, "
                                           f"{T} steps, dt 0.1", n=n, span=1000,
                                  xi0_rule=f"c/(n maxd), c={c}", xi0=xi0, valid=int(len(ok)),
                                  valid_fraction=len(ok) / A,
                                  best=float(ok.min()) if len(ok) else None,
                                  mean=float(ok.mean()) if len(ok) else None,
                                  random_tour=t.random_length(a.seed),
                                  steps_executed=info["steps_executed"], **tsp_energies(s, lens)))
                s.close()
    elif a.workload == "quality":
        # Best and mean cut of the narrow-state forms against plain FP32 dSB
        # at equal steps on several graph families (N = 2^18, 256 agents, dt
        # 0.5, 20,000 steps, no window exit), three seeds each.  The bar is a
        # best cut within 0.1% of plain FP32; the line states which pass.
        T = a.steps
        fams = [("4-regular +-1", lambda g: sis.Csr.gen_regular(1 << 18, 4, sis.W_PM1, g)),
                ("3-regular +-1", lambda g: sis.Csr.gen_regular(1 << 18, 3, sis.W_PM1, g)),
                ("3-regular +1", lambda g: sis.Csr.gen_regular(1 << 18, 3, sis.W_ONE, g)),
                ("4-regular +1", lambda g: sis.Csr.gen_regular(1 << 18, 4, sis.W_ONE, g)),
                ("ER mean degree 8 +-1", lambda g: sis.Csr.gen_er(1 << 18, 8.0 / (1 << 18),
                                                                  sis.W_PM1, g)),
                ("ER mean degree 8, weights {-3..3}\\{0}",
                 lambda g: sis.Csr.gen_er(1 << 18, 8.0 / (1 << 18), sis.W_INT3, g)),
                ("torus 512x512 +-1", lambda g: sis.Csr.gen_torus(512, 512, sis.W_PM1, g))]
        for fam, mk in fams:
            for g in (a.seed, a.seed + 1, a.seed + 2):
                J = mk(g)
                res = {}
                for form in ("f32", "q12", "q10", "q8", "q8s"):
                    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=256, dt=0.5,
                                           form=sis.FORMS[form]))
                    s.init_random(g + 100)
                    info = s.run(T, 0)
                    c, _ = s.cuts()
                    res[form] = dict(best=float(c.max()), mean=float(c.mean()),
                                     ms_per_step=info["step_kernel_ms"] / info["steps_executed"])
                    s.close()
                b = res["f32"]
                for f in ("q12", "q10", "q8", "q8s"):
                    r = res[f]
                    r["best_rel_diff"] = (r["best"] - b["best"]) / abs(b["best"])
                    r["mean_rel_diff"] = (r["mean"] - b["mean"]) / abs(b["mean"])
                    r["within_0.1pct"] = bool(r["best_rel_diff"] >= -1e-3)
                # printed as each comparison ends (the whole sweep takes minutes)
                print(json.dumps(dict(workload=f"quality: {fam}, N={J.n}, 256 agents, dt 0.5, "
                                               f"{T} steps, dSB, no window exit", seed=g,
                                      nnz=J.nnz, plain=b, q12=res["q12"], q10=res["q10"],
                                      q8=res["q8"], q8s=res["q8s"])),
                      flush=True)
    elif a.workload == "powerlaw":
        # Chung-Lu N = 2^20, mean degree 8, exponent 2.1 (largest row ~1.2e5
        # entries), +-1 weights, 256 agents, dSB dt 0.5; without and with the
        # row split (threshold chosen by the autotuner).  dSB with integer
        # weights is bit-identical either way, so the cuts must agree.
        J = sis.Csr.gen_chung_lu(1 << 20, 8.0, 2.1, sis.W_PM1, a.seed)
        rp, _, _ = J.arrays()
        T = a.steps if a.steps != 20000 else 2000
        for form in ("packed", "f32"):
            for split in (0, -1):
                r = run_solver(J, dict(problem=sis.MAXCUT, agents=256, dt=0.5, form=sis.FORMS[form],
                                       split=split), T, 0, a.seed)
                ms = r["info"]["step_kernel_ms"] / max(r["info"]["steps_executed"], 1)
                nbytes = step_bytes(form, J.n, J.nnz, 256)
                lines.append(dict(workload=f"powerlaw Chung-Lu N=2^20 avg 8 gamma 2.1, +-1, dSB, 256 "
                                           f"agents, {T} steps", form=form, nnz=J.nnz,
                                  max_row=int(np.diff(rp).max()), split_requested=split,
                                  **{k: v for k, v in r.items() if k != "info"},
                                  step_kernel_ms_per_step=ms,
                                  achieved_GBs=nbytes / (ms * 1e-3) / 1e9,
                                  ms_per_step=r["wall"] * 1e3 / max(r["info"]["steps_executed"], 1)))
    for ln in lines:
        print(json.dumps(ln), flush=True)


def batch_stream(a):
    """The batch engine's streaming step alone (the fallback of resident
    chunks): the serving mix (N log-uniform in [64, 4096], ER mean degree 8,
    weights {-3..3}\\{0}, 64 agents, dt 0.5) admitted whole, then `steps`
    steps of each tile in --batch-tiles, direct launches, one JSON line each."""
    rng = np.random.default_rng(a.seed)
    A, count = 64, a.instances
    sizes = np.floor(np.exp(rng.uniform(np.log(64), np.log(4096), count))).astype(int)
    Js = [sis.Csr.gen_er(int(n), 8.0 / (n - 1), sis.W_INT3, 7000 + k) for k, n in enumerate(sizes)]
    rows, nnz = int(sizes.sum()), int(sum(J.nnz for J in Js))
    K = a.steps if a.steps != 20000 else 300
    for tile in [int(t, 0) for t in a.batch_tiles.split(",") if t]:
        b = sis.Batch(agents_max=A, max_instances=count, max_rows=rows, max_nnz=nnz,
                      chunk_steps=100, tile=tile, eager=True)
        for k, J in enumerate(Js):
            b.submit(J, A, 0.5, 1 << 30, 0, seed=k)
        b.advance(0)
        b.advance(100)  # warm-up chunk
        t0 = time.perf_counter()
        b.advance(K)
        wall = time.perf_counter() - t0
        b.close()
        state = rows * A * 16  # X and Y read and written, FP32
        print(json.dumps(dict(workload=f"batch streaming step, {count} MaxCut instances "
                                       f"({rows} rows, {nnz} stored entries), 64 agents",
                              tile=hex(tile), steps=K, ms_per_step=wall * 1e3 / K,
                              state_bytes=state, state_GBs=state / (wall / K) / 1e9)), flush=True)


def serving_run(seed, count, T, chunk, eager, rate, seq_sample):
    """Batched serving of many small MaxCut dSB instances: N log-uniform in
    [64, 4096], Erdos-Renyi with mean degree 8, weights uniform in
    {-3..3}\\{0}, 64 agents, dt 0.5, T steps, update window on (instances
    retire when every agent has locked).  Closed loop (everything submitted at
    once), then Poisson arrivals, then a sample solved one at a time with the
    plain FP32 solver.  The plain FP32 dSB trajectory is bit-identical to the
    packed one the batch runs, so each sampled instance's cuts and spins must
    equal its result in the closed-loop mix."""
    rng = np.random.default_rng(seed)
    A, dt = 64, 0.5
    sizes = np.floor(np.exp(rng.uniform(np.log(64), np.log(4096), count))).astype(int)
    Js = [sis.Csr.gen_er(int(n), 8.0 / (n - 1), sis.W_INT3, 7000 + k) for k, n in enumerate(sizes)]
    rows, nnz = int(sizes.sum()), int(sum(J.nnz for J in Js))

    def engine(tile=0):
        return sis.Batch(agents_max=A, max_instances=count, max_rows=rows, max_nnz=nnz,
                         chunk_steps=chunk, tile=tile, eager=eager)

    # Tile choice on the full resident set, outside the timed regions.
    b = engine()
    for k, J in enumerate(Js):
        b.submit(J, A, dt, T, 1, seed=k)
    b.advance(0)
    tile, _ = b.autotune(10)
    b.close()

    # Closed loop: everything submitted at t = 0.
    b = engine(tile)
    t0 = time.perf_counter()
    tickets = [b.submit(J, A, dt, T, 1, seed=k) for k, J in enumerate(Js)]
    while True:
        running, queued = b.advance(chunk)
        if running == 0 and queued == 0:
            break
    wall = time.perf_counter() - t0
    st = [b.poll(t) for t in tickets]
    steps_done = sum(x["steps_executed"] for x in st)
    closed = dict(solves_per_s=count / wall, wall_s=wall, mean_steps=steps_done / count,
                  node_agent_steps_per_s=sum(x["steps_executed"] * x["n"] * A for x in st) / wall)
    sample = list(range(0, count, max(1, count // seq_sample)))[:seq_sample]
    mixed = {k: b.result(tickets[k], spins=True) for k in sample}
    mixed_steps = {k: st[k]["steps_executed"] for k in sample}
    b.close()

    # Poisson arrivals at rate lam (default: 70% of the closed-loop rate).
    lam = rate if rate > 0 else 0.7 * closed["solves_per_s"]
    arrivals = np.cumsum(rng.exponential(1.0 / lam, count))
    b = engine(tile)
    start = time.perf_counter()
    nxt, delay = 0, {}
    while True:
        now = time.perf_counter() - start
        while nxt < count and arrivals[nxt] <= now:
            t = b.submit(Js[nxt], A, dt, T, 1, seed=nxt)
            delay[t] = (time.perf_counter() - start) - arrivals[nxt]  # arrival -> submit
            nxt += 1
        running, queued = b.advance(chunk)
        if nxt == count and running == 0 and queued == 0:
            break
        if running == 0 and queued == 0:
            time.sleep(max(0.0, arrivals[nxt] - (time.perf_counter() - start)))
    done = {t: b.poll(t) for t in delay}
    # latency = scheduled arrival -> observed retirement (library clock for
    # submit -> done, plus the loop's arrival -> submit delay)
    lats = np.array([x["done_s"] - x["submit_s"] + delay[t] for t, x in done.items()])
    span = max(x["done_s"] for x in done.values()) - min(x["submit_s"] for x in done.values())
    poisson = dict(rate_per_s=lam, solves_per_s=count / span, p50_s=float(np.percentile(lats, 50)),
                   p99_s=float(np.percentile(lats, 99)), mean_s=float(lats.mean()))
    b.close()

    # Sequential plain FP32 solves of the sample (graph-replayed window
    # periods); the time covers create + init + run + cuts of each solve.
    seq_time, same = 0.0, 0
    for k in sample:
        t0 = time.perf_counter()
        s = sis.Solver(Js[k], dict(problem=sis.MAXCUT, agents=A, dt=dt, form=sis.FORM_F32,
                                   use_graph=0 if eager else 1))
        s.init_random(k)
        info = s.run(T, 1)
        cuts, _ = s.cuts()
        seq_time += time.perf_counter() - t0
        mc, ms = mixed[k]
        same += int(np.array_equal(cuts, mc) and np.array_equal(s.spins(), ms)
                    and info["steps_executed"] == mixed_steps[k])
        s.close()
    seq = len(sample) / seq_time
    return dict(count=count, rows=rows, nnz=nnz, tile=tile, closed=closed, poisson=poisson,
                seq=seq, sample=len(sample), same=same)


def serving(a):
    r = serving_run(a.seed, a.instances, a.steps, a.chunk, a.eager, a.rate, a.seq_sample)
    out = {
        "metric": "solves/s (batched serving) vs sequential plain FP32 solves",
        "value": r["closed"]["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": 0,
        "higher_is_better": True,
        "dtype": "f32",
        "data": "synthetic (seeded Erdos-Renyi instances, seeded initial state)",
        "config": {"workload": f"{r['count']} MaxCut instances, N log-uniform [64,4096] (sum "
                               f"{r['rows']}), ER mean degree 8, weights {{-3..3}}\\{{0}}, 64 "
                               f"agents, dt 0.5, {a.steps} steps, update window on, chunk "
                               f"{a.chunk} steps",
                   "autotuned_tile": r["tile"], "nnz": r["nnz"]},
        "closed_loop": r["closed"],
        "poisson": r["poisson"],
        "sequential_plain_solves_per_s": r["seq"],
        "sequential_sample": r["sample"],
        "speedup_vs_sequential_plain": r["closed"]["solves_per_s"] / r["seq"],
        "alone_vs_mix_identical": f"{r['same']}/{r['sample']}",
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--form", default="q8s", choices=list(sis.FORMS))
    ap.add_argument("--baseline-form", default="f32", choices=list(sis.FORMS) + ["none"])
    ap.add_argument("--nodes-log2", type=int, default=20)
    ap.add_argument("--agents", type=int, default=256)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-bf16", action="store_true", help="skip the BF16 bSB legs")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving sub-object")
    ap.add_argument("--no-graph", action="store_true",
                    help="skip the whole-run hipGraph replay of the headline solve")
    ap.add_argument("--also", default="q10,q12,q8,packed", help="comma-separated extra forms "
                    "timed beside the headline (same dynamics); the default adds the 10-, 12- "
                    "and 8-bit (nearest-even) narrow-state forms and the packed form "
                    "(bit-identical to plain FP32)")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--workload", default="regular2p20",
                    choices=["regular2p20", "er800", "torus100", "tsp48", "powerlaw", "tsp_struct",
                             "tsp_valid", "serving", "batch_stream", "quality"],
                    help="regular2p20: the headline (random 4-regular graph, N = 2^20); the "
                         "others print one JSON line per configuration")
    ap.add_argument("--instances", type=int, default=4096)
    ap.add_argument("--chunk", type=int, default=100)
    ap.add_argument("--rate", type=float, default=0.0)
    ap.add_argument("--seq-sample", type=int, default=32)
    ap.add_argument("--eager", action="store_true",
                    help="serving: launch chunk kernels directly instead of replaying a hipGraph")
    ap.add_argument("--batch-tiles", default="0x90004,0x2090004,0x2090008",
                    help="batch_stream: comma-separated streaming tiles of the batch engine")
    a = ap.parse_args()
    a.also = [f for f in a.also.split(",") if f]
    if a.gpus != 1:
        print("bench.py measures one GPU: the solver runs on a single device (agents shard "
              "across GPUs with no per-step communication); use --gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    if a.workload == "serving":
        if a.steps == 20000:
            a.steps = 5000
        serving(a)
        return
    if a.workload == "batch_stream":
        batch_stream(a)
        return
    if a.workload in ("er800", "torus100", "tsp48", "powerlaw", "tsp_struct", "tsp_valid",
                      "quality"):
        small_configs(a)
        return

    N = 1 << a.nodes_log2
    A, dt, K = a.agents, 0.5, a.steps
    J = sis.Csr.gen_regular(N, 4, sis.W_PM1, a.seed)
    Z = J.nnz
    xi0 = J.xi0()

    def leg(form, dyn):
        r = time_form(J, form, A, dt, K, a.warmup, a.seed, dyn)
        info = r["info"]
        ex = max(info["steps_executed"], 1)
        step_ms = info["step_kernel_ms"] / ex
        nbytes = step_bytes(form, N, Z, A)
        r.update(form=form, executed=info["steps_executed"], value=info["steps_executed"] * A / r["wall"],
                 ms_per_step=r["wall"] * 1e3 / ex, avg_launch_ms=step_ms, bytes=nbytes,
                 achieved=nbytes / (step_ms * 1e-3) / 1e9)
        return r

    def summary(r):
        return {"form": r["form"], "value": r["value"], "ms_per_step": r["ms_per_step"],
                "avg_launch_ms": r["avg_launch_ms"], "achieved_GBs": r["achieved"],
                "frac": r["achieved"] / HBM_PEAK_GBS, "best_cut": r["best_cut"],
                "mean_cut": r["mean_cut"], "autotuned_tile": r["tune"],
                "energy_check": r["energy_check"]}

    def cut_diff(r, base):
        return (r["best_cut"] - base["best_cut"]) / abs(base["best_cut"])

    # The BF16 form is bSB; its plain baseline runs bSB too.
    dyn = sis.BSB if a.form == "bf16" else sis.DSB
    head = leg(a.form, dyn)
    base = None
    if a.baseline_form != "none" and a.baseline_form != a.form:
        base = leg(a.baseline_form, dyn)
    executed = head["executed"]
    traffic = measured_traffic(a.form, (head["tune"] or {}).get("tile"), N, Z, A)
    out = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "agent-steps/s",
        "n_gpus": 1,
        "steps": executed,
        "warmup": a.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if a.form != "f64" else "f64",
        "state_storage": {"q12": "X, Y as 12-bit fixed-point codes (3 bytes per node-agent)",
                          "q10": "X, Y as 10-bit fixed-point codes (2.5 bytes per node-agent)",
                          "q8": "X, Y as 8-bit fixed-point codes (2 bytes per node-agent)",
                          "q8s": "X, Y as 8-bit fixed-point codes (2 bytes per node-agent), "
                                 "Y rounded stochastically",
                          "packed": "X, Y FP32", "f32": "X, Y FP32", "bf16": "X, Y FP32",
                          "f64": "X, Y FP64"}[a.form],
        "data": "synthetic (seeded random 4-regular graph, +-1 weights, seeded initial state)",
        "config": {
            "workload": f"maxcut {'bSB' if dyn == sis.BSB else 'dSB'}, random 4-regular "
                        f"N=2^{a.nodes_log2}, Z={Z}, +-1 weights, {A} agents, dt {dt}, "
                        f"{executed} steps executed ({K} requested; a run executes whole "
                        f"window periods of 100 steps), update window every 100 steps "
                        f"(locks counted, early exit off)",
            "form": a.form,
            "nodes": N,
            "nnz": Z,
            "agents": A,
            "xi0": xi0,
            "steps_requested": K,
            "autotuned_tile": head["tune"],
        },
        "roofline": {
            "bound": "hbm",
            "achieved": head["achieved"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": head["achieved"] / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": ("profiles/pmc_traffic_maxcut_4regular_N2p20_A256.json: rocprofv3 "
                               "FETCH_SIZE + WRITE_SIZE of this kernel tile, calibrated (not "
                               "measured in this run)") if traffic is not None else None,
            "kernel": f"step ({a.form})",
            "bytes_per_launch": head["bytes"],
            "avg_launch_ms": head["avg_launch_ms"],
            "timing": "HIP events around the step launches on the solver's stream",
        },
        "device_ms": head["info"]["device_ms"],
        "step_kernel_ms": head["info"]["step_kernel_ms"],
        "best_cut": head["best_cut"],
        "best_agent": head["best_agent"],
        "mean_cut": head["mean_cut"],
        "energy_check": head["energy_check"],
    }
    if base is not None:
        out["plain"] = summary(base)
        out["speedup_vs_plain"] = head["value"] / base["value"]
        out["best_cut_rel_diff_vs_plain"] = cut_diff(head, base)
    for form in a.also:
        if form in (a.form, a.baseline_form):
            continue
        r = leg(form, dyn)
        d = summary(r)
        if base is not None:
            d["speedup_vs_plain"] = r["value"] / base["value"]
            d["best_cut_rel_diff_vs_plain"] = cut_diff(r, base)
        out[form] = d
    if not a.no_graph:
        out["graph"] = graph_run(J, a.form, A, dt, K, a.seed, dyn, (head["tune"] or {}).get("tile"),
                                 head)
    if not a.no_bf16 and dyn == sis.DSB:
        # bSB with BF16 gathers against plain FP32 bSB at equal steps
        rb, rp = leg("bf16", sis.BSB), leg("f32", sis.BSB)
        eh, eb = rb["cont_energy"], rp["cont_energy"]
        out["bf16_bsb"] = dict(
            summary(rb), plain_bsb=summary(rp), speedup_vs_plain_bsb=rb["value"] / rp["value"],
            best_cut_rel_diff_vs_plain_bsb=cut_diff(rb, rp),
            continuous_energy_vs_plain_bsb=dict(
                agents=len(eb), max_rel_diff=float(np.max(np.abs(eh - eb) / np.abs(eb))),
                mean_rel_diff=float(np.mean(np.abs(eh - eb) / np.abs(eb))),
                host="scipy FP64 (1/2) x^T J x of each run's final FP32 state"))
    if not a.no_serving:
        r = serving_run(a.seed, a.instances, 5000, a.chunk, a.eager, a.rate, a.seq_sample)
        out["serving"] = dict(
            workload=f"{r['count']} MaxCut instances, N log-uniform [64,4096], ER mean degree 8, "
                     f"weights {{-3..3}}\\{{0}}, 64 agents, dt 0.5, 5000 steps, update window on",
            closed_loop_solves_per_s=r["closed"]["solves_per_s"],
            poisson_rate_per_s=r["poisson"]["rate_per_s"],
            poisson_solves_per_s=r["poisson"]["solves_per_s"],
            p50_s=r["poisson"]["p50_s"], p99_s=r["poisson"]["p99_s"],
            sequential_plain_solves_per_s=r["seq"],
            speedup_vs_sequential_plain=r["closed"]["solves_per_s"] / r["seq"],
            alone_vs_mix_identical=f"{r['same']}/{r['sample']}", tile=r["tile"])
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(J, A, dt, xi0, a.cpu_seconds, a.seed)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

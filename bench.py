"""Benchmark: FBPINN train steps/s (and collocation points/s) of the fused,
graph-captured step on one MI355X, against the plain per-op GPU form of the
same method and a bounded CPU sample of the oracle restatement.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--config wave]
python bench.py --config C --train-l1     (full-length training, test L1 history)

Prints ONE JSON line.  Default workload: the (2+1)D wave configuration
(1,048,576 collocation points, 12x12x6 subdomains, overlap 1.9, [3,64,64,1]).
Default K and W per configuration (DEFAULT_STEPS): about a second of timed
steps (the wave's 100 steps, 20,000 of the oscillator's 25 us steps), so that
the first replay's start-up is not a visible share of the timed region.
"""

import argparse
import json
import os
import re
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# (timed steps, warm-up steps) when --steps / --warmup are not given
DEFAULT_STEPS = {"wave": (100, 10), "poisson": (2000, 100), "burgers": (5000, 200), "ho": (20000, 1000),
                 "inverse": (10000, 500)}

PEAK_FP32_MFMA = 157.3   # TFLOP/s, MI355X dense FP32 matrix
PEAK_HBM = 8000.0         # GB/s


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def measured_traffic(config, kernel, launch):
    """HBM bytes per launch of `kernel` on `config`, as recorded from the
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes in profiles/<config>_traffic.md
    (a line `<kernel> traffic_bytes_per_launch = N launch = <shape>`), or None
    when no figure was recorded for this run's launch shape `launch`."""
    path = os.path.join(ROOT, "profiles", f"{config}_traffic.md")
    if not os.path.exists(path):
        return None
    pat = re.compile(rf"^\s*{re.escape(kernel)}\s+traffic_bytes_per_launch\s*=\s*(\d+)\s+launch\s*=\s*(.+?)\s*$")
    with open(path) as f:
        for line in f:
            m = pat.match(line)
            if m and m.group(2) == launch:
                return int(m.group(1))
    return None


def parse_launch(text):
    """'bins2,chunks512' -> (("bins", 2), ("chunks", 512)): forward, backward
    (wavesC: C-pair chunks, one wave each, for one-hidden-layer networks)."""
    specs = []
    for part in text.split(","):
        m = re.fullmatch(r"(bins|chunks|waves)(\d+)", part.strip())
        if not m:
            raise SystemExit(f"--launch: {part!r} is not binsK, chunksC or wavesC")
        specs.append((m.group(1), int(m.group(2))))
    if len(specs) != 2:
        raise SystemExit("--launch takes two shapes: forward,backward")
    return tuple(specs)


def calibrate_copy(mib):
    """A plain copy kernel of known size (reads and writes `mib` MiB each
    launch): the reference for scaling FETCH_SIZE / WRITE_SIZE counters."""
    n = mib * (1 << 20) // 4
    src = torch.rand(n, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"copy_kernel_bytes_read": n * 4, "copy_kernel_bytes_written": n * 4, "launches": 5}))


def train_l1(name):
    """Train configuration `name` for its full step count through
    FBPINNTrainer (trainers.py:577-687) and print one JSON line: test L1 and
    normalised L1 against the problem's exact solution at every test step
    (trainers.py:727-774), the steady-state rate, and for the batched inverse
    configuration the learnt damping coefficients against the true ones."""
    from fbpinns_amd import configs
    from fbpinns_amd.trainers import FBPINNTrainer
    from fbpinns_amd.util.logger import logger
    for h in logger.handlers:
        h.setStream(sys.stderr)
    c = configs.CONFIGS[name]()
    tr = FBPINNTrainer(c, write_outputs=False)
    t0 = time.time()
    ap = tr.train()
    rec = {"config": name, "n_steps": c.n_steps, "wall_s": round(time.time() - t0, 2),
           "test_points": int(np.prod(c.n_test)),
           "l1": [{"step": int(r[0]), "l1": r[4], "l1_normalised": r[5]} for r in tr.u_test_losses]}
    if "mu" in ap["trainable"].get("problem", {}):
        mu = ap["trainable"]["problem"]["mu"].detach().double().cpu().numpy().reshape(-1)
        mu_true = np.asarray(ap["static"]["problem"]["mu_true"], dtype=np.float64).reshape(-1)
        rel = np.abs(mu - mu_true) / np.abs(mu_true)
        w0 = np.asarray(ap["static"]["problem"]["w0"], dtype=np.float64).reshape(-1)
        worst = np.argsort(-rel)[:8]
        rec["mu_rel_err"] = {"median": float(np.median(rel)), "mean": float(rel.mean()), "max": float(rel.max()),
                             "instances": int(rel.size), "within_5pct": int((rel < 0.05).sum()),
                             "worst": [{"instance": int(i), "w0": float(w0[i]), "mu_true": float(mu_true[i]),
                                        "mu": float(mu[i])} for i in worst]}
    print(json.dumps(rec), flush=True)


def matched_l1(name, steps, point_chunks, over):
    """The fused engine and the plain per-op form from identical parameters and
    points, `steps` train steps each under the scheduler's first state, with
    the test L1 of both against the problem's exact solution (trainers.py:
    727-774) at four points along the way; prints one JSON line."""
    from fbpinns_amd import configs
    from fbpinns_amd.plain import PlainFBPINN
    from fbpinns_amd.trainers import FBPINNTrainer
    from fbpinns_amd.util.logger import logger
    for h in logger.handlers:
        h.setStream(sys.stderr)
    c = configs.CONFIGS[name](**over)
    tr = FBPINNTrainer(c, write_outputs=False)
    eng, sched, x_test, u_exact = tr.setup(exact=True)
    active = next(iter(sched))
    pl = PlainFBPINN(c.problem, c.decomposition, c.network, eng,
                     point_chunks=point_chunks if len(eng.cglobal) == 1 else 1)
    pl.set_active(active)
    eng.set_active(active)
    xt = x_test.to(eng.device, torch.float32).contiguous()
    ue = u_exact.to(eng.device).reshape(-1, 1)
    sd = ue.std().item()

    def l1(u):
        return torch.mean(torch.abs(ue - u.reshape(-1, 1))).item()
    marks = sorted({max(1, steps * k // 4) for k in range(1, 5)})
    hist, t_f, t_p, done = [], 0.0, 0.0, 0
    for mk in marks:
        torch.cuda.synchronize()
        t0 = time.time()
        for _ in range(mk - done):
            eng.step()
        torch.cuda.synchronize()
        t_f += time.time() - t0
        t0 = time.time()
        t_log = t0
        for j in range(mk - done):
            pl.step()
            if time.time() - t_log >= 30.0:   # a line per half minute of plain steps
                torch.cuda.synchronize()
                t_log = time.time()
                log(f"plain step {done + j + 1} of {steps}")
        torch.cuda.synchronize()
        t_p += time.time() - t0
        done = mk
        a, b = l1(eng.model(xt)), l1(pl.model(xt))
        hist.append({"step": mk, "l1_fused": a, "l1_plain": b, "rel_diff": abs(a - b) / b})
        log(f"step {mk}: test L1 fused {a:.6e} plain {b:.6e} (rel diff {abs(a - b) / b:.2e})")
    rec = {"config": name, "steps": steps, "test_points": int(xt.shape[0]), "l1_fused": hist[-1]["l1_fused"],
           "l1_plain": hist[-1]["l1_plain"], "l1n_fused": hist[-1]["l1_fused"] / sd,
           "l1n_plain": hist[-1]["l1_plain"] / sd, "rel_diff": hist[-1]["rel_diff"], "history": hist,
           "fused_steps_per_s": steps / t_f, "plain_steps_per_s": steps / t_p,
           "speedup": t_p / t_f, "plain_point_chunks": pl.point_chunks}
    print(json.dumps(rec), flush=True)


def time_kernel(fn, reps):
    """Average duration (ms) of fn() launched on torch's current stream,
    measured with HIP events on that stream."""
    s = torch.cuda.current_stream()
    fn()
    s.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(s)
    for _ in range(reps):
        fn()
    b.record(s)
    b.synchronize()
    return a.elapsed_time(b) / reps


def kernel_roofline(eng, cfgmod):
    from fbpinns_amd import ops
    plan, net = eng.plan, eng.net
    wl = plan.wl
    # fill the backward inputs once (gseg) with a real step's gradient
    loss, g, _ = eng.loss_and_grad()
    torch.cuda.synchronize()

    def fwd():
        ops.forward(net.as_list(), eng.theta, eng.dom, plan.x, wl.tensor_list(), wl.size_list(), plan.pairbuf)

    def bwd():
        ops.backward(net.as_list(), eng.theta, eng.dom, plan.x, wl.tensor_list(), wl.size_list(),
                     plan.chunk_train, plan.gseg, plan.partial)

    t_fwd = time_kernel(fwd, 10)
    t_bwd = time_kernel(bwd, 10)
    f_pair = cfgmod.fwd_flops_per_pair(net.xd, net.h, net.nh, net.J)
    trained_pairs = int((plan.chunk_train.cpu().numpy().astype(bool)
                         * np.diff(wl.t["chunk_start"].cpu().numpy())).sum())
    fl_fwd = f_pair * wl.n_pairs
    fl_bwd = 2 * f_pair * trained_pairs
    # what the matrix cores execute (the kernels' algebra runs fewer
    # contractions than the plain jet count above)
    n1, n2 = bin(net.d1mask).count("1"), bin(net.d2mask).count("1")
    x_fwd, x_bwd = cfgmod.mfma_flops_per_pair(net.h, net.nh, n1, n2)
    ex_fwd, ex_bwd = x_fwd * wl.n_pairs, x_bwd * trained_pairs
    tf = lambda fl, t: fl / (t * 1e-3) / 1e12
    return dict(t_fwd_ms=t_fwd, t_bwd_ms=t_bwd, flop_fwd=fl_fwd, flop_bwd=fl_bwd,
                tflops_fwd=tf(fl_fwd, t_fwd), tflops_bwd=tf(fl_bwd, t_bwd),
                f_pair=f_pair, n_pairs=wl.n_pairs, trained_pairs=trained_pairs,
                executed_mfma={"flop_per_pair_fwd": x_fwd, "flop_per_pair_bwd": x_bwd,
                               "tflops_fwd": tf(ex_fwd, t_fwd), "tflops_bwd": tf(ex_bwd, t_bwd),
                               "frac_fwd": tf(ex_fwd, t_fwd) / PEAK_FP32_MFMA,
                               "frac_bwd": tf(ex_bwd, t_bwd) / PEAK_FP32_MFMA})


def host_threads():
    """CPUs this process may use: its affinity mask, capped by the cgroup CPU
    quota when one is set (a container's share of a larger host)."""
    n = len(os.sched_getaffinity(0))
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            quota, period = f.read().split()[:2]
        if quota != "max":
            n = min(n, max(1, int(int(quota) // int(period))))
    except (OSError, ValueError):
        pass
    return n


def cpu_baseline(c, eng, budget_s=20.0, min_steps=10, max_steps=1000):
    """The oracle (torch on CPU, float32) on a bounded sample of the same
    points -- every k-th of the (Z-ordered) training points, so the sample
    spans the domain -- all subdomains, on every CPU this process may use;
    times as many steps as fit the budget (min_steps to max_steps: 1,000 for
    the oscillator's full 200 points, 10-13 on the wave's sample) and
    returns steps/s extrapolated by pair count."""
    from oracle import adam as oadam
    from oracle import model as om
    from oracle import problems as opb
    from oracle import takes as ot
    torch.set_num_threads(host_threads())
    cores = torch.get_num_threads()
    ap = eng.all_params
    p = ap["static"]["decomposition"]["subdomain"]["params"]
    xmins, xmaxs = p[0].cpu().numpy(), p[1].cpu().numpy()
    pou = ap["static"]["decomposition"]["subdomain"]["pou"].cpu().numpy().reshape(-1)
    prob = ap["static"]["problem"]
    name = c.problem.__name__
    if name == "WaveEquationConstantVelocity3D":
        req, C, L = opb.wave(prob["source"].cpu().numpy(), prob["c0"])
    elif name == "PoissonMultiscale2D":
        req, C, L = opb.poisson(prob["sd"], prob["ks"])
    elif name == "BurgersEquation2D":
        req, C, L = opb.burgers(prob["nu"], prob["sd"])
    elif name == "HarmonicOscillator1DHardBC":
        req, C, L = opb.ho_hard(prob["d"], prob["w0"], prob["sd"])
    else:
        return None
    xall = eng.x.cpu().numpy()
    n_full = xall.shape[0]
    n = min(n_full, 1024)
    while True:
        stride = max(1, n_full // n)
        x = xall[::stride][:n]
        takes, all_ims, *_ = ot.get_inputs(x, np.ones(len(pou), dtype=int), xmins, xmaxs, pou)
        layers = [(W.detach().cpu()[all_ims].float(), b.detach().cpu()[all_ims].float())
                  for W, b in ap["trainable"]["network"]["subdomain"]["layers"]]
        dom = om.to_torch_dom(xmins[all_ims], xmaxs[all_ims], p[4].cpu().numpy()[all_ims],
                              p[5].cpu().numpy()[all_ims], torch.float32)
        jm = [om.get_jmaps(r) for r in req]
        cons = [[torch.tensor(x)]]

        def one_step(layers):
            loss, g, _ = om.value_and_grad(layers, dom, cons, [takes], jm, C, L)
            out = []
            for (W, b), (gW, gb) in zip(layers, g):
                Wn, *_ = oadam.adam_update(W.numpy(), np.zeros_like(W.numpy()), np.zeros_like(W.numpy()),
                                           gW.numpy(), 0)
                bn, *_ = oadam.adam_update(b.numpy(), np.zeros_like(b.numpy()), np.zeros_like(b.numpy()),
                                           gb.numpy(), 0)
                out.append((torch.tensor(Wn), torch.tensor(bn)))
            return out
        t0 = time.time()
        layers = one_step(layers)
        t1 = time.time() - t0
        # grow the sample until min_steps steps take about the budget
        if t1 * min_steps * 1.5 > budget_s or n == n_full:
            break
        n = min(n_full, int(n * max(1.2, min(8.0, budget_s / (min_steps * max(t1, 1e-3))))))
    # the count from a warm step (the first call above includes set-up)
    t0 = time.time()
    layers = one_step(layers)
    t1 = time.time() - t0
    steps = max(min_steps, min(max_steps, int(budget_s / max(t1, 1e-4))))
    t0 = time.time()
    for _ in range(steps):
        layers = one_step(layers)
    dt = (time.time() - t0) / steps
    s_sample = len(takes[0])
    full_pairs = eng.plan.wl.n_pairs
    rate = (s_sample / dt) / full_pairs
    return {"value": rate, "unit": "steps/s", "cores": cores, "kind": "port", "cpu_model": _cpu_model(),
            "sample": f"oracle (torch CPU float32, nested jvp + autograd + Adam) on every {stride}th of the "
                      f"{n_full} Z-ordered points ({x.shape[0]} points, {s_sample} of {full_pairs} pairs), "
                      f"{steps} timed steps of {dt:.2f} s on {cores} threads; steps/s extrapolated by pair count"}


def _cpu_model():
    """lscpu's model name and topology, with the CPUs this process may use."""
    info = {}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    name = info.get("Model name")
    if not name:
        try:
            with open("/proc/cpuinfo") as f:
                name = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "unknown")
        except OSError:
            name = "unknown"
    topo = ", ".join(f"{k} {info[k]}" for k in ("CPU(s)", "Socket(s)", "Core(s) per socket", "Thread(s) per core")
                     if k in info)
    return f"{name} ({topo}; {host_threads()} usable by this process)"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default: the configuration's)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps first (default: the configuration's)")
    ap.add_argument("--config", default="wave")
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--plain-chunks", type=int, default=16)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-graph", action="store_true", help="launch the step eagerly (for counter passes)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shape (counter passes see one launch shape)")
    ap.add_argument("--launch", default=None, metavar="F,B",
                    help="fix the forward,backward launch shapes (binsK / chunksC / wavesC, e.g. bins2,bins2) "
                         "instead of autotuning them")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain copy kernel of MIB MiB (counter calibration) and exit")
    ap.add_argument("--overlap", type=float, default=None,
                    help="wave/poisson/burgers: subdomain overlap factor in place of the configuration's")
    ap.add_argument("--train-l1", action="store_true",
                    help="only train --config for its full step count and print its test L1 history")
    ap.add_argument("--matched-l1", type=int, default=0, metavar="K",
                    help="only train the fused and the plain forms side by side for K steps from the same "
                         "start and print both test L1 errors")
    a = ap.parse_args()
    k_default, w_default = DEFAULT_STEPS.get(a.config, (100, 10))
    a.steps = k_default if a.steps is None else a.steps
    a.warmup = w_default if a.warmup is None else a.warmup
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (every run here uses one MI355X)",
              file=sys.stderr)
        sys.exit(2)
    launch = parse_launch(a.launch) if a.launch else None
    if a.calibrate_copy:
        calibrate_copy(a.calibrate_copy)
        return
    if a.train_l1:
        train_l1(a.config)
        return

    from fbpinns_amd import autotune, configs
    from fbpinns_amd.plain import PlainFBPINN
    from fbpinns_amd.trainers import FBPINNTrainer
    from fbpinns_amd.util.logger import logger
    for h in logger.handlers:
        h.setStream(sys.stderr)   # stdout carries only the JSON line

    over = {}
    if a.overlap is not None:
        if a.config not in ("wave", "poisson", "burgers"):
            print(f"--overlap applies to the wave, poisson and burgers configurations", file=sys.stderr)
            sys.exit(2)
        xs = configs.CONFIGS[a.config]().decomposition_init_kwargs["subdomain_xs"]
        lo, hi = [v[0] for v in xs], [v[-1] for v in xs]
        over = dict(decomposition_init_kwargs=configs._rect(lo, hi, [len(v) for v in xs], a.overlap))
    if a.matched_l1:
        matched_l1(a.config, a.matched_l1, a.plain_chunks, over)
        return
    c = configs.CONFIGS[a.config](graph=not a.no_graph, autotune=not a.no_autotune, **over)
    log(f"config {a.config}: building problem")
    tr = FBPINNTrainer(c, write_outputs=False)
    eng, sched, x_test, _ = tr.setup(exact=False)
    active = next(iter(sched))
    t0 = time.time()
    eng.set_active(active, launch=launch)
    torch.cuda.synchronize()
    log(f"work list + graph capture {time.time() - t0:.1f} s: {eng.plan.wl.n_points} points, "
        f"{eng.plan.wl.n_pairs} pairs")
    for _ in range(a.warmup):
        eng.step()
    torch.cuda.synchronize()
    log("timing fused steps")
    t0 = time.time()
    for _ in range(a.steps):
        loss = eng.step()
    torch.cuda.synchronize()
    dt = time.time() - t0
    sps = a.steps / dt
    n_pts = eng.plan.wl.n_points
    log(f"fused: {sps:.2f} steps/s, loss {loss.item():.4e}")

    rl = kernel_roofline(eng, configs)
    log(f"kernels: fwd {rl['t_fwd_ms']:.3f} ms ({rl['tflops_fwd']:.1f} TFLOP/s), "
        f"bwd {rl['t_bwd_ms']:.3f} ms ({rl['tflops_bwd']:.1f} TFLOP/s)")
    dom_is_bwd = rl["t_bwd_ms"] >= rl["t_fwd_ms"]
    ach = rl["tflops_bwd"] if dom_is_bwd else rl["tflops_fwd"]
    dom_kernel = "bwd_kernel" if dom_is_bwd else "fwd_kernel"
    # the recorded counters hold for the configuration's own overlap and the
    # launch shape they were measured with
    wgf, wgb = eng.plan.wl.workgroups()
    launch_txt = (f"fwd {autotune.name(eng.launch[0])} ({wgf} workgroups), "
                  f"bwd {autotune.name(eng.launch[1])} ({wgb} workgroups)")
    traffic = measured_traffic(a.config, dom_kernel, launch_txt) if not over else None

    plain = None
    if not a.no_plain:
        log("plain per-op form")
        eng2 = FBPINNTrainer(configs.CONFIGS[a.config](graph=False, **over),
                             write_outputs=False).setup(exact=False)[0]
        pl = PlainFBPINN(c.problem, c.decomposition, c.network, eng2,
                         point_chunks=a.plain_chunks if len(eng2.cglobal) == 1 else 1)
        pl.set_active(active)
        eng2.set_active(active)
        l_plain0 = pl.step().item()
        l_fused0 = eng2.step().item()
        torch.cuda.synchronize()
        t0 = time.time()
        for _ in range(a.plain_steps):
            pl.step()
        torch.cuda.synchronize()
        pdt = (time.time() - t0) / a.plain_steps
        # matched solutions: the fused step from the same start, as many
        # steps, then both models on the test grid
        for _ in range(a.plain_steps):
            eng2.step()
        xt = x_test.to(eng2.device, torch.float32).contiguous()
        u_f, u_p = eng2.model(xt), pl.model(xt)
        u_diff = ((u_f - u_p).abs().mean() / u_p.abs().mean()).item()
        plain = {"steps_per_s": 1.0 / pdt, "ms_per_step": pdt * 1e3, "point_chunks": pl.point_chunks,
                 "first_step_loss_rel_diff_vs_fused": abs(l_plain0 - l_fused0) / abs(l_fused0),
                 "matched_after_steps": 1 + a.plain_steps, "test_points": int(xt.shape[0]),
                 "test_u_rel_l1_diff_vs_fused": u_diff}
        log(f"plain: {1.0 / pdt:.4f} steps/s; speed-up {sps * pdt:.1f}x; after {1 + a.plain_steps} steps "
            f"mean |u_fused - u_plain| / mean |u_plain| = {u_diff:.2e} on {xt.shape[0]} test points")
        del pl, eng2
        torch.cuda.empty_cache()

    cpu = None
    if not a.no_cpu:
        log("cpu baseline (oracle sample)")
        cpu = cpu_baseline(c, eng, a.cpu_budget)

    out = {
        "metric": "FBPINN train steps/s + collocation pts/s at matched L1 error, fused vs plain",
        "value": sps,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 / sps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded collocation points, random-init subdomain networks)",
        "config": {"workload": f"{a.config}{'' if not over else f' (overlap {a.overlap})'}: "
                               f"{n_pts} collocation points, {eng.m} subdomains, "
                               f"{eng.plan.wl.n_pairs} (point, subdomain) pairs, FCN [{eng.net.xd}, "
                               f"{eng.net.h}x{eng.net.nh}, 1], jets J={eng.net.J}, hipGraph-captured step",
                   "collocation_points": n_pts, "subdomains": eng.m, "pairs": eng.plan.wl.n_pairs,
                   "launch": {"forward": f"{autotune.name(eng.launch[0])}, {wgf} workgroups",
                              "backward": f"{autotune.name(eng.launch[1])}, {wgb} workgroups"}},
        "collocation_pts_per_s": sps * n_pts,
        "roofline": {"bound": "mfma", "achieved": ach, "peak": PEAK_FP32_MFMA, "unit": "TFLOP/s",
                     "frac": ach / PEAK_FP32_MFMA, "traffic": traffic,
                     "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated, "
                                     "recorded for this launch shape)",
                     "kernel": dom_kernel, "launch": launch_txt,
                     "executed_frac": rl["executed_mfma"]["frac_bwd" if dom_is_bwd else "frac_fwd"]},
        "kernels": rl,
        "plain_gpu": plain,
        "speedup_vs_plain": (sps / plain["steps_per_s"]) if plain else None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""Training loop for the trainable decoder on the MI355X.

Mirrors deep_learning/ldpc_model_trainer.py (``train_c_table_ldpc_model``
:49-323) and deep_learning/ldpc_training_data.py
(``make_c_table_training_batch`` :24-82): BCE-with-logits (or softplus
margin) loss on -LLR of the B payload+CRC bits, optional intermediate
supervision every ``intermediate_loss_stride`` iterations, float64 global
gradient-norm clipping to 1.0, Adam with cosine annealing to 0.05 * lr, and
the best-loss state kept on the host.  The unrolled decode and its gradient
are the library's HIP kernels (see ``unrolled.py``); training inputs are
synthesized on the GPU (``ldpc_mi355_synth_device``: payload, CRC, encoding,
rate matching, channel and the float rate recovery of the reference's
training data), or on host threads with ``on_device=False``.

New here: ``per_codeblock_noise`` draws one noise power per codeblock from
``training_noise_power_db_range`` (the reference draws one per batch).

    python -m deeplearning_5g_ldpc_decoder_amd.trainer --steps 20 --batch 256
"""
import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

from .unrolled import (CTableNeuralLayeredMinSumDecoder, load_ldpc_table_config,
                       synth_float_batch, synth_float_batch_device)

CRC_TYPES = (0, 1, 2, 3)


def make_c_table_training_batch(model, batch_size, noise_power_db, device, seed=0,
                                threads=16, on_device=True):
    """(llr [batch, n] float32, target [batch, B] float32) on ``device``.
    noise_power_db: a scalar or one value per codeblock.  on_device: synthesize
    on the GPU (ldpc_mi355_synth_device); otherwise on host threads."""
    import torch
    nz = np.broadcast_to(np.asarray(noise_power_db, np.float32), (batch_size,))
    if on_device:
        return synth_float_batch_device(model.config, nz, seed=seed)
    llr, bits = synth_float_batch(model.config, nz, seed=seed, threads=threads)
    return (torch.from_numpy(llr).to(device, non_blocking=True),
            torch.from_numpy(bits.astype(np.float32)).to(device, non_blocking=True))


def _grad_norm_float64(params):
    total = 0.0
    for p in params:
        if p.grad is None:
            continue
        g = p.grad.detach()
        if not bool(g.isfinite().all()):
            return float("inf")
        total += float((g.double() * g.double()).sum().cpu())
    return total ** 0.5


def _clip_grad_norm_float64(params, max_norm):
    params = [p for p in params if p.grad is not None]
    norm = _grad_norm_float64(params)
    if not np.isfinite(norm):
        return norm
    coef = max_norm / (norm + 1e-6)
    if coef < 1.0:
        for p in params:
            p.grad.detach().mul_(coef)
    return norm


def train_c_table_ldpc_model(bgn, A, crctype, E, rv, Qm, noise_power_db, HARQ_on=0,
                             num_iterations=8, epochs=50, batch_size=8, lr=1e-2, init_alpha=0.7,
                             init_beta=0.0, learnable_scaling=True, learnable_beta=False,
                             alpha_mode="single", alpha_min=0.5, alpha_max=1.0 - 1e-6,
                             init_alpha_degree_values=None, clip_llr=None,
                             detach_check_update_inputs=False,
                             training_noise_power_db_range=None, loss_type="bce",
                             llr_margin=5.0, device=None, cpu_threads=4,
                             save_path="trained_c_table_ldpc_decoder.pth", lib_path=None,
                             intermediate_loss_weight=0.0, intermediate_loss_stride=4,
                             verbose=True, beta_mode="checkpoint", per_codeblock_noise=False,
                             seed=0, synth_threads=16, history=None):
    """Train alpha/beta; returns (model at its best step, best state dict,
    best loss).  ``epochs`` counts optimizer steps, one fresh batch each, as
    in the reference.  ``history`` (a list) receives one dict per step."""
    import torch
    if crctype not in CRC_TYPES:
        raise ValueError("crctype must be 0, 1, 2, or 3")
    if intermediate_loss_weight < 0:
        raise ValueError("intermediate_loss_weight must be non-negative")
    if intermediate_loss_stride <= 0:
        raise ValueError("intermediate_loss_stride must be positive")
    if loss_type not in {"bce", "margin"}:
        raise ValueError("loss_type must be 'bce' or 'margin'")
    if llr_margin <= 0:
        raise ValueError("llr_margin must be positive")
    if training_noise_power_db_range is not None:
        lo, hi = training_noise_power_db_range
        if lo > hi:
            raise ValueError("training_noise_power_db_range must be in (low, high) order")
    if cpu_threads is not None:
        torch.set_num_threads(cpu_threads)
    device = torch.device(device) if device is not None else torch.device("cuda")
    if device.type != "cuda":
        raise RuntimeError("the trainable decoder runs on the GPU only")
    table_config = load_ldpc_table_config(bgn, A, crctype, E, rv, Qm, HARQ_on, lib_path)
    model = CTableNeuralLayeredMinSumDecoder(
        table_config, num_iterations=num_iterations, learnable_scaling=learnable_scaling,
        learnable_beta=learnable_beta, init_alpha=init_alpha, init_beta=init_beta,
        alpha_mode=alpha_mode, alpha_min=alpha_min, alpha_max=alpha_max,
        init_alpha_degree_values=init_alpha_degree_values, clip_llr=clip_llr,
        detach_check_update_inputs=detach_check_update_inputs, beta_mode=beta_mode).to(device)
    params = [p for p in model.parameters() if p.requires_grad]
    if not params:
        raise ValueError("nothing to train: learnable_scaling and learnable_beta are both off")
    optimizer = torch.optim.Adam(params, lr=lr)
    scheduler = torch.optim.lr_scheduler.CosineAnnealingLR(optimizer, T_max=epochs,
                                                           eta_min=lr * 0.05)
    criterion = torch.nn.BCEWithLogitsLoss()
    B = table_config.B

    def decoder_loss(llr, target):
        if loss_type == "bce":
            return criterion(-llr[:, :B], target)
        signed = (1.0 - 2.0 * target) * llr[:, :B]
        return torch.nn.functional.softplus(llr_margin - signed).mean()

    rng = np.random.default_rng(seed)
    best_loss = float("inf")
    best_state = {k: v.detach().cpu().clone() for k, v in model.state_dict().items()}
    if verbose:
        mem = model.memory_summary(batch_size)
        print(f"training C-table LDPC model: bgn={bgn} A={A} B={B} Zc={table_config.Zc} "
              f"n={model.n} rows={model.num_rows} edges={model.num_edges} "
              f"alpha_mode={alpha_mode} message_memory={mem['total_message_MB']:.2f} MB",
              flush=True)
    for epoch in range(epochs):
        if training_noise_power_db_range is not None:
            lo, hi = training_noise_power_db_range
            nz = rng.uniform(lo, hi, batch_size if per_codeblock_noise else 1)
        else:
            nz = np.array([noise_power_db], np.float64)
        llr, target = make_c_table_training_batch(
            model, batch_size, nz if nz.size == batch_size else float(nz[0]), device,
            seed=seed * 1000003 + epoch, threads=synth_threads)
        model.train()
        optimizer.zero_grad(set_to_none=True)
        t0 = time.time()
        want_it = intermediate_loss_weight > 0
        out = model(llr, return_llr=True, return_iteration_llrs=want_it)
        final_llr = out[1]
        if not bool(torch.isfinite(final_llr).all()):
            raise FloatingPointError("non-finite final_llr detected before loss; try a lower "
                                     "alpha_max, clip_llr, or fewer iterations")
        final_loss = decoder_loss(final_llr, target)
        inter = final_loss.new_zeros(())
        if want_it:
            # every stride-th iteration's beliefs, the last one included when
            # the stride divides the iteration count: the reference's
            # per-iteration outputs are clones, never the final tensor itself
            # (ldpc_model_trainer.py:207-210)
            sup = out[2][intermediate_loss_stride - 1::intermediate_loss_stride]
            if sup:
                inter = torch.stack([decoder_loss(x, target) for x in sup]).mean()
            loss = final_loss + intermediate_loss_weight * inter
        else:
            loss = final_loss
        if not bool(torch.isfinite(loss)):
            raise FloatingPointError("non-finite training loss detected")
        loss.backward()
        norm = _clip_grad_norm_float64(model.parameters(), 1.0)
        if not np.isfinite(norm):
            raise FloatingPointError("non-finite model gradient detected")
        optimizer.step()
        scheduler.step()
        lv = float(loss.item())
        dt = time.time() - t0
        if history is not None:
            history.append(dict(step=epoch, loss=lv, final_loss=float(final_loss.item()),
                                grad_norm=norm, seconds=dt))
        if lv < best_loss:
            best_loss = lv
            best_state = {k: v.detach().cpu().clone() for k, v in model.state_dict().items()}
        if verbose and (epoch == 0 or (epoch + 1) % 5 == 0 or epoch + 1 == epochs):
            a = model.alpha().detach()
            print(f"epoch {epoch + 1}/{epochs} loss={lv:.6f} "
                  f"final_loss={float(final_loss.item()):.6f} "
                  f"intermediate_loss={float(inter.item()):.6f} "
                  f"alpha_mean={float(a.mean()):.4f} alpha_min={float(a.min()):.6f} "
                  f"alpha_max={float(a.max()):.6f} grad_norm={norm:.3e} "
                  f"beta_mean={float(model.beta().detach().mean()):.4f} "
                  f"lr={optimizer.param_groups[0]['lr']:.3e} step_s={dt:.4f}", flush=True)
    model.load_state_dict(best_state)
    if save_path is not None:
        Path(save_path).parent.mkdir(parents=True, exist_ok=True)
        torch.save({"state_dict": best_state, "table_config": table_config.__dict__,
                    "best_loss": best_loss}, save_path)
    return model, best_state, best_loss


# The trained-decoder workload: BG1 Zc=96 at rate 1/3, QPSK, Es/N0 uniform in
# [0, 2] dB per codeblock, 10 unrolled iterations, alpha and beta per
# (iteration, row degree) initialised at 0.7 / 0.0, batch 2048, Adam 1e-3.
BG1_ZC96_WORKLOAD = dict(bgn=1, A=2088, crctype=2, E=6336, rv=0, Qm=2, num_iterations=10,
                         alpha_mode="iteration_degree", beta_mode="iteration_degree",
                         init_alpha=0.7, init_beta=0.0, learnable_scaling=True,
                         learnable_beta=True, lr=1e-3, batch_size=2048,
                         training_noise_power_db_range=(-2.0, 0.0), per_codeblock_noise=True)


def evaluate_algo2_vs_algo1(model, noise_power_dbs, n=4096, seed=0, alpha_q15=22936):
    """Block error rates of the model's exported algo-2 tables, installed in
    the library (load_algo2_tables), against the int16 min-sum with one
    alpha (algo 1), both for the model's number of iterations with CRC early
    stop, on the same synthesized codeblocks (n per noise power).  The
    library's default algo-2 tables are restored afterwards."""
    import torch
    from .batch import BatchPlan, synth_batch
    from .tables import set_trained_tables
    from .unrolled import load_algo2_tables
    c = model.config
    cfg = (c.bgn, c.A, c.crctype, c.E, c.rv, c.Qm, 0)
    T = model.num_iterations
    rows = {"algo2_trained": [2, T, 0, 0, 1280, 32767, -32767],
            "algo1": [1, T, alpha_q15, 0, 1280, 32767, -32767]}
    load_algo2_tables(model)
    out = []
    try:
        for k, nz in enumerate(noise_power_dbs):
            llr, _ = synth_batch(cfg, n, float(nz), seed=seed * 1000 + k, want_payload=False)
            d = torch.from_numpy(llr.reshape(-1)).cuda()
            point = {"noise_power_db": float(nz), "EsN0_dB": -float(nz), "codeblocks": n}
            for name, row in rows.items():
                plan = BatchPlan([cfg] * n, row)
                res = plan.decode(d)
                st = res["status"].cpu().numpy()
                point[name] = {"bler": float(np.mean(st != 0)),
                               "mean_iterations": float(res["iters"].float().mean())}
                plan.close()
            out.append(point)
    finally:
        set_trained_tables(c.bgn)
    return {"decoder_rows": rows, "points": out}


def main(argv=None):
    ap = argparse.ArgumentParser(
        description="Train alpha/beta of the BG1 Zc=96 workload (rate 1/3 QPSK, 10 unrolled "
                    "iterations, per-(iteration, degree) parameters, batch 2048, Adam 1e-3 with "
                    "cosine annealing) and export the algo-2 tables.")
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--batch", type=int, default=2048)
    ap.add_argument("--noise-range", type=float, nargs=2, default=None,
                    metavar=("LOW", "HIGH"),
                    help="noise power range in dB, one draw per codeblock (default -2 0: "
                         "Es/N0 uniform in [0, 2] dB)")
    ap.add_argument("--out", default=None, help="directory for the algo-2 table export")
    ap.add_argument("--save", default=None)
    ap.add_argument("--history", default=None)
    ap.add_argument("--evaluate", type=float, nargs="*", default=None, metavar="NOISE_DB",
                    help="after training: BLER of the exported algo-2 tables against algo 1 "
                         "at these noise powers (dB)")
    ap.add_argument("--summary", default=None, help="JSON summary of the run")
    args = ap.parse_args(argv)
    from .unrolled import algo2_tables, export_algo2_tables
    cfg = dict(BG1_ZC96_WORKLOAD)
    if args.noise_range is not None:
        cfg["training_noise_power_db_range"] = tuple(args.noise_range)
    cfg.update(epochs=args.steps, batch_size=args.batch, save_path=args.save, noise_power_db=-1.0,
               verbose=False)
    hist = []
    t0 = time.time()
    model, _, best = train_c_table_ldpc_model(history=hist, **cfg)
    wall = time.time() - t0
    print(f"{args.steps} steps in {wall:.1f} s, best loss {best:.6g}", flush=True)
    summary = {"workload": {k: (list(v) if isinstance(v, tuple) else v)
                            for k, v in cfg.items() if k != "save_path"},
               "steps": args.steps, "wall_s": wall,
               "loss": {"first": hist[0]["loss"], "last": hist[-1]["loss"], "best": best,
                        "every_100": [h["loss"] for h in hist[::100]]}}
    if args.out:
        summary["exported"] = [str(p) for p in export_algo2_tables(model, args.out)]
        print("exported", summary["exported"], flush=True)
    fa, fb = algo2_tables(model, num_iterations=model.num_iterations)
    summary["algo2_alpha_q15"] = fa.tolist()
    summary["algo2_beta_llr128"] = fb.tolist()
    if args.evaluate:
        summary["evaluation"] = evaluate_algo2_vs_algo1(model, args.evaluate)
        for p in summary["evaluation"]["points"]:
            print(f"noise {p['noise_power_db']:+.2f} dB: algo 2 (trained) BLER "
                  f"{p['algo2_trained']['bler']:.4f}, algo 1 BLER {p['algo1']['bler']:.4f}",
                  flush=True)
    if args.history:
        with open(args.history, "w") as f:
            json.dump(hist, f)
    if args.summary:
        with open(args.summary, "w") as f:
            json.dump(summary, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""Torch FP32 restatement of the reference trainable decoder's forward
(deep_learning/deeplearning_ldpc_model.py:315-410), used in tests as the
autograd reference for the oracle's hand-derived gradient and the HIP
kernels: the same tensor operations (gather, where, abs, min with argmin,
one-hot mask, masked min, product of signs, clamp, scatter into a clone of
LQ), so torch autograd produces the reference's gradient conventions.
Test infrastructure only.
"""
import torch


def lifted_columns(table_config):
    """[row] -> LongTensor [Zc * degree] of lifted column indices (:168-184)."""
    zc = table_config.Zc
    zx = torch.arange(zc, dtype=torch.long).unsqueeze(1)
    out = []
    ro = table_config.row_offsets
    for r in range(table_config.num_rows):
        cols = torch.tensor(table_config.hbg_columns[ro[r]:ro[r + 1]], dtype=torch.long)
        sh = torch.tensor(table_config.vij[ro[r]:ro[r + 1]], dtype=torch.long)
        out.append((cols.unsqueeze(0) * zc + (sh.unsqueeze(0) + zx) % zc).reshape(-1))
    return out


def unrolled_forward(table_config, llr, alpha, beta, clip=None, detach=False):
    """llr [batch, n]; alpha/beta [T, rows] tensors (may require grad).
    Returns (final LQ, list of per-iteration LQ)."""
    zc = table_config.Zc
    cols = lifted_columns(table_config)
    degrees = [len(c) // zc for c in cols]
    batch = llr.shape[0]
    LQ = llr.clone()
    if clip is not None:
        LQ = LQ.clamp(-clip, clip)
    Lr = [llr.new_zeros((batch, zc, d)) for d in degrees]
    iters = []
    for t in range(alpha.shape[0]):
        for r, deg in enumerate(degrees):
            rel = LQ[:, cols[r]].reshape(batch, zc, deg) - Lr[r]
            upd = rel.detach() if detach else rel
            signs = torch.where(upd < 0, -torch.ones_like(upd), torch.ones_like(upd))
            a = upd.abs()
            m1, i1 = torch.min(a, dim=2)
            mask = torch.nn.functional.one_hot(i1, num_classes=deg).to(torch.bool)
            m2, _ = torch.min(a.masked_fill(mask, float("inf")), dim=2)
            sprod = torch.prod(signs, dim=2, keepdim=True)
            use = torch.where(mask, m2.unsqueeze(2), m1.unsqueeze(2))
            mag = torch.clamp(use - beta[t, r], min=0.0)
            mag = alpha[t, r] * mag
            new_lr = sprod * signs * mag
            q = upd + new_lr
            if clip is not None:
                q = q.clamp(-clip, clip)
            nxt = LQ.clone()
            nxt[:, cols[r]] = q.reshape(batch, -1)
            LQ = nxt
            Lr[r] = new_lr
        iters.append(LQ.clone())
    return LQ, iters

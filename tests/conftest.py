import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _split(arr, lens):
    out, o = [], 0
    for n in lens:
        out.append(arr[o:o + n])
        o += n
    return out


def load_decode_cases(name):
    """Golden decode cases written by tests/golden/gen_golden.py."""
    z = np.load(os.path.join(GOLDEN, name))
    llr = _split(z["llr"], z["llr_len"])
    hard = _split(z["hard"], z["hard_len"])
    soft = _split(z["soft"], z["buf_len"])
    lq = _split(z["lq"], z["buf_len"])
    harq = _split(z["harq_in"], z["harq_len"])
    cases = []
    for i in range(len(z["cfg"])):
        c = dict(idx=i, cfg=tuple(int(v) for v in z["cfg"][i]),
                 row=[int(v) for v in z["row"][i]],
                 status=int(z["status"][i]), iters=int(z["iters"][i]),
                 llr=llr[i], hard=hard[i], soft=soft[i], lq=lq[i],
                 harq_in=harq[i] if len(harq[i]) else None)
        if "table" in z.files:
            # algo 2: the (alpha, beta) [32, rows] tables the reference build had
            t, bgn = int(z["table"][i]), c["cfg"][0]
            c["table_id"] = t
            c["tables"] = (z[f"alpha_bg{bgn}"][t], z[f"beta_bg{bgn}"][t])
        if "degree_table" in z.files:
            # algo 3: None = the reference defaults through c_ldpc_5g_decode
            t = int(z["degree_table"][i])
            c["table_id"] = t
            c["degree_alpha"] = None if t < 0 else z["degree_alpha"][t]
        cases.append(c)
    return cases


def case_id(c):
    bgn, A, crc, E, rv, Qm, h = c["cfg"]
    tab = f"-tab{c['table_id']}" if "table_id" in c else ""
    return (f"{c['idx']}-bg{bgn}-A{A}-E{E}-rv{rv}-Qm{Qm}-h{h}-algo{c['row'][0]}"
            f"-it{c['row'][1]}{tab}")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle as O
    O.lib()
    return O


def rate_recovery_cases():
    """Reference HARQ_memory fixtures after rate recovery
    (tests/golden/gen_golden.py: repetition with int16 wrap, an rv chain with
    HARQ combining): (cfg7, filler, llr, soft_before or None, soft_after)."""
    z = np.load(os.path.join(GOLDEN, "rate_recovery.npz"))
    from oracle import oracle as O
    out = []
    lo = so = 0
    prev = None
    for m in z["meta"]:
        bgn, A, crctype, E, rv, Qm, harq_on, filler = (int(v) for v in m)
        st, p = O.params(bgn, A, crctype, E, rv, Qm, harq_on)
        n = p["N"] + 2 * p["Zc"]
        llr = z["llr"][lo:lo + E]
        lo += E
        want = z["soft"][so:so + n]
        so += n
        out.append(((bgn, A, crctype, E, rv, Qm, harq_on), filler, llr,
                    prev if harq_on else None, want))
        prev = want
    return out

"""One hgcd! level on the device (src/gcd.jl:546-554): hgcd_matrix_adjust!
(src/gcd.jl:398-437) then hgcd_matrix_mul! (src/gcd.jl:372-391) on resident
buffers updated in place, for 1024 instances (n0 log-uniform in [600, 32768]),
against Python big ints with libgmp products.  A second run on the updated
buffers checks that the level is sized for changing contents."""
import sys

import numpy as np
import pytest

from tests import gmp_ref
from tests.helpers import wl

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
from nbhip import device as D  # noqa: E402
from nbhip import engine  # noqa: E402

SENT = np.uint64(0xA5A5A5A5A5A5A5A5)
BETA = 1 << 64


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _padded(x, cap):
    out = np.full(cap + 2, SENT, dtype=np.uint64)  # two sentinel limbs past the capacity
    out[:x.size] = x
    out[x.size:cap] = 0
    return out


def _normlen(v):
    return (v.bit_length() + 63) // 64


def _expected(q, sh, a_limbs, b_limbs, M_limbs):
    """Signed adjust results and the matrix product of one instance, from the
    host limbs the level reads: the low p limbs of a, b and the low mn limbs
    of the M entries (libgmp products, Python-int sums)."""
    p, mn = sh["p"], sh["mn"]
    al, bl = a_limbs[:p], b_limbs[:p]
    M = {k: M_limbs[k][:mn] for k in M_limbs}
    pairs = [(q["n11"], al), (q["n01"], bl), (q["n00"], bl), (q["n10"], al),
             (M["m00"], q["n00"]), (M["m01"], q["n10"]), (M["m00"], q["n01"]), (M["m01"], q["n11"]),
             (M["m10"], q["n00"]), (M["m11"], q["n10"]), (M["m10"], q["n01"]), (M["m11"], q["n11"])]
    return pairs


def _combine(q, sh, prods):
    P = [wl.limbs_to_int(x) for x in prods]
    shift = BETA ** sh["p"]
    ea = wl.limbs_to_int(q["ra"]) * shift + P[0] - P[1]
    eb = wl.limbs_to_int(q["rb"]) * shift + P[2] - P[3]
    eM = {"m00": P[4] + P[5], "m01": P[6] + P[7], "m10": P[8] + P[9], "m11": P[10] + P[11]}
    return ea, eb, eM


@pytest.mark.parametrize("count", [1024])
def test_level_in_place_two_runs(count):
    engine.init()
    assert gmp_ref.available()
    sizes = wl.hgcd_sizes(count)
    host, dev, shapes = [], [], []
    for k, n0 in enumerate(sizes):
        q = wl.hgcd_level_instance(k, n0)
        sh = wl.hgcd_level_shape(n0)
        d = {}
        for name in ("m00", "m01", "m10", "m11"):
            d[name] = D.to_dev(_padded(q[name], sh["Lm"]))
        for name in ("n00", "n01", "n10", "n11", "ra", "rb"):
            d[name] = D.to_dev(q[name])
        for name in ("a", "b"):
            # the values' n limbs, and room for the L-limb results
            d[name] = D.to_dev(_padded(q[name], max(q[name].size, sh["L"])))
        d["mn"], d["p"] = sh["mn"], sh["p"]
        host.append(q)
        dev.append(d)
        shapes.append(sh)
    say(f"hgcd level: {count} instances, n0 up to {max(sizes)}")
    H = D.HgcdLevel(dev)
    try:
        a_now = [D.to_host(d["a"]) for d in dev]
        b_now = [D.to_host(d["b"]) for d in dev]
        M_now = [{k: q[k] for k in ("m00", "m01", "m10", "m11")} for q in host]
        for rep in range(2):
            pairs = []
            for i, (q, sh) in enumerate(zip(host, shapes)):
                pairs += _expected(q, sh, a_now[i], b_now[i], M_now[i])
            prods = gmp_ref.products(pairs)
            H.run()
            torch.cuda.synchronize()
            signs = H.signs.cpu().numpy()
            lens = H.lens.cpu().numpy()
            bad = []
            for i, (q, d, sh) in enumerate(zip(host, dev, shapes)):
                ea, eb, eM = _combine(q, sh, prods[12 * i:12 * i + 12])
                ga, gb = D.to_host(d["a"]), D.to_host(d["b"])
                L, Lm = sh["L"], sh["Lm"]
                # limbs from L on are left as they were (the reference rewrites a, b through L only)
                ok = (int(signs[i, 0]) == (ea > 0) - (ea < 0) and int(signs[i, 1]) == (eb > 0) - (eb < 0) and
                      wl.limbs_to_int(ga[:L]) == abs(ea) and wl.limbs_to_int(gb[:L]) == abs(eb) and
                      np.array_equal(ga[L:], a_now[i][L:]) and np.array_equal(gb[L:], b_now[i][L:]) and
                      int(lens[i, 0]) == max(_normlen(abs(ea)), _normlen(abs(eb)), 1))
                gm = {}
                for k in ("m00", "m01", "m10", "m11"):
                    x = D.to_host(d[k])
                    gm[k] = x[:Lm]
                    ok = ok and wl.limbs_to_int(gm[k]) == eM[k] and bool((x[Lm:] == SENT).all())
                ok = ok and int(lens[i, 1]) == max(max(_normlen(v) for v in eM.values()), 1)
                if not ok:
                    bad.append(i)
                a_now[i], b_now[i], M_now[i] = ga, gb, gm
            say(f"run {rep}: {count - len(bad)}/{count} levels exact")
            assert not bad, bad[:10]
    finally:
        H.close()

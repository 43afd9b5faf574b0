"""Host-side plan tables of the register FFT engine (csrc/r2r.h, csrc/fft_reg.h), checked on
the CPU: every compiled radix sequence multiplies to its length, its values per thread are a
multiple of every radix, and the LDS exchange positions of each pass (fft_reg.h xpos, restated
here) form a permutation that the next pass reads back, inside the per-line stride
seq_lds_stride gives (a stride too small for the radix-4 passes was the failure this guards)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = open(os.path.join(ROOT, "hos-nwt_amd", "csrc", "r2r.h")).read()


def table(fn):
    body = SRC[SRC.index(f"inline std::vector<int> {fn}(int L, bool alt)"):]
    body = body[:body.index("default:")]
    out = {}
    for m in re.finditer(r"case (\d+): return (alt \? std::vector<int>\{([\d, ]+)\} : std::vector<int>\{([\d, ]+)\}"
                         r"|\{([\d, ]+)\});", body):
        L = int(m.group(1))
        if m.group(3):
            out[L] = ([int(v) for v in m.group(4).split(",")], [int(v) for v in m.group(3).split(",")])
        else:
            seq = [int(v) for v in m.group(5).split(",")]
            out[L] = (seq, seq)
    return out


def values(fn):
    body = SRC[SRC.index(f"inline int {fn}(int L)"):]
    body = body[:body.index("\n}\n")]
    out = {}
    for m in re.finditer(r"((?:case \d+: )+)return (\d+);", body):
        for L in re.findall(r"case (\d+):", m.group(1)):
            out[int(L)] = int(m.group(2))
    d = re.search(r"default: return (\d+);", body)
    return out, int(d.group(1))


def padx(p):
    return p + (p >> 3)


def xpos(yp, R, Ns, p, W):
    if yp:
        return p + (p >> 3)
    if Ns == 1:
        return p + p // W
    if Ns < W:
        return p + Ns * (p // (R * Ns))
    return p


def lds_stride(L, rs, eb):  # seq_lds_stride
    W = 256 // eb
    mx = padx(L - 1)
    Ns = 1
    for R in rs[:-1]:
        p = L - 1
        x = p + p // W if Ns == 1 else (p + Ns * (p // (R * Ns)) if Ns < W else p)
        mx = max(mx, x)
        Ns *= R
    lp = mx + 1
    return lp + (lp & 1)


SEQ = table("seq_radices")
V0, V0d = values("seq_values")
V1, _ = values("seq_alt_values")


def test_tables_parse():
    assert set(SEQ) == {12288, 8192, 6144, 3072, 1536, 768, 4096, 2048, 1024, 512}
    assert V1[6144] == 12 and V0[6144] == 24 and V1[12288] == 12


@pytest.mark.parametrize("L", sorted(SEQ))
@pytest.mark.parametrize("alt", [False, True])
def test_sequence_exchanges_fit_the_stride(L, alt):
    rs = SEQ[L][1] if alt and L in V1 else SEQ[L][0]
    V = V1[L] if alt and L in V1 else V0.get(L, V0d)
    prod = 1
    for R in rs:
        prod *= R
        assert V % R == 0, (L, rs, V)
    assert prod == L
    ntl = L // V
    assert ntl <= 1024  # threads of one line (one workgroup)
    for eb in (8, 4):
        W = 256 // eb
        for yp in (False, True):
            lp = lds_stride(L, rs, eb)
            Ns = 1
            for i, R in enumerate(rs[:-1]):
                R2 = rs[i + 1]
                wpos = {}
                for tl in range(ntl):
                    for b in range(V // R):
                        j = tl + b * ntl
                        k = j % Ns
                        o = (j - k) * R + k
                        for r in range(R):
                            e = o + r * Ns
                            q = xpos(yp, R, Ns, e, W)
                            assert q not in wpos
                            wpos[q] = e
                assert max(wpos) < lp, (L, rs, eb, yp, max(wpos), lp)
                for tl in range(ntl):
                    for b in range(V // R2):
                        j = tl + b * ntl
                        for r in range(R2):
                            e = j + r * (L // R2)
                            assert wpos.get(xpos(yp, R, Ns, e, W)) == e
                Ns *= R

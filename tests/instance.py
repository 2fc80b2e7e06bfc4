"""Loading a FORMAT.md key directory (bend2/FORMAT.md:15-24) for the CPU oracle, and the
closed-form Groth16 proof from the setup trapdoor (trapdoor.bin).

Closed form (libsnark reduction, ref/README.md:24-33): with A(t), B(t), C(t) the QAP
polynomials interpolating the evaluation vectors (A with the input-consistency rows),
IC = sum_{i < num_instance} z_i (beta u_i + alpha v_i + w_i)(t):
  a = alpha + A(t) + r delta,   b = beta + B(t) + s delta,
  c = (beta A(t) + alpha B(t) - IC + A(t) B(t)) / delta + s a + r b - r s delta,
and the Groth16 verification equation in the exponent reads a b = alpha beta + IC + c delta.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

from oracle import oracle as O

R = O.P_R


def _words(path: Path) -> np.ndarray:
    return np.fromfile(path, dtype=np.uint32)


def _pack(words16: np.ndarray, per: int) -> np.ndarray:
    """16-bit-limb words -> (n, per) uint64 (per = 4 for one field element)."""
    w = words16.astype(np.uint64).reshape(-1, 4)
    v = w[:, 0] | (w[:, 1] << np.uint64(16)) | (w[:, 2] << np.uint64(32)) | (w[:, 3] << np.uint64(48))
    return np.ascontiguousarray(v.reshape(-1, per))


def _csr(path: Path, nc: int):
    w = _words(path)
    nnz = int(w[0])
    assert w.size == 2 + nc + 17 * nnz, f"{path}: bad size"
    off = np.ascontiguousarray(w[1:2 + nc])
    ent = w[2 + nc:].reshape(nnz, 17) if nnz else np.zeros((0, 17), np.uint32)
    col = np.ascontiguousarray(ent[:, 0])
    coef = _pack(np.ascontiguousarray(ent[:, 1:]).reshape(-1), 4)
    return off, col, coef


def load(d: str | Path) -> dict:
    d = Path(d)
    hdr = _words(d / "header.bin")
    nc, ni, nw, lg, na, nb, nh, nl = (int(x) for x in hdr)
    inst = {"header": hdr, "dir": d}
    for m in "abc":
        inst[f"{m}_off"], inst[f"{m}_col"], inst[f"{m}_coef"] = _csr(d / f"r1cs_{m}.bin", nc)
    inst["z"] = _pack(_words(d / "witness.bin"), 4)
    inst["pk_g1"] = _pack(_words(d / "pk_g1.bin"), 8)
    inst["pk_g2"] = _pack(_words(d / "pk_g2.bin"), 16)
    for name, per in (("a", 8), ("b1", 8), ("b2", 16), ("h", 8), ("l", 8)):
        inst[f"pk_{name}"] = _pack(_words(d / f"pk_{name}.bin"), per)
    rs = _pack(_words(d / "rs.bin"), 4)
    inst["r"], inst["s"] = np.ascontiguousarray(rs[0]), np.ascontiguousarray(rs[1])
    if (d / "trapdoor.bin").exists():
        td = _pack(_words(d / "trapdoor.bin"), 4)
        inst["trapdoor"] = {k: O.unmont(O.from_words(td[i]), R)
                            for i, k in enumerate(("tau", "alpha", "beta", "gamma", "delta"))}
    return inst


def _row_evals(off, col, coef, z: list[int], nc: int) -> list[int]:
    coef_i = [O.unmont(O.from_words(c), R) for c in coef]
    out = []
    for i in range(nc):
        acc = 0
        for e in range(int(off[i]), int(off[i + 1])):
            acc += coef_i[e] * z[int(col[e])]
        out.append(acc % R)
    return out


def eval_vectors(inst: dict) -> tuple[list[int], list[int], list[int]]:
    nc, ni, nw, lg = (int(x) for x in inst["header"][:4])
    d = 1 << lg
    z = [O.unmont(O.from_words(v), R) for v in inst["z"]]
    a = _row_evals(inst["a_off"], inst["a_col"], inst["a_coef"], z, nc) + z[:ni]
    b = _row_evals(inst["b_off"], inst["b_col"], inst["b_coef"], z, nc)
    c = _row_evals(inst["c_off"], inst["c_col"], inst["c_coef"], z, nc)
    pad = lambda v: v + [0] * (d - len(v))
    return pad(a), pad(b), pad(c)


def eval_vectors_mont(inst: dict) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """eval_vectors as (d, 4) u64 Montgomery arrays through the oracle's C rows (the
    same values; test_oracle_math pins the two against each other)."""
    nc, ni, nw, lg = (int(x) for x in inst["header"][:4])
    d = 1 << lg
    out = []
    for m in "abc":
        v = O.eval_rows(inst[f"{m}_off"], inst[f"{m}_col"], inst[f"{m}_coef"], inst["z"], nc, d)
        if m == "a":
            v[nc:nc + ni] = inst["z"][:ni]
        out.append(v)
    return out[0], out[1], out[2]


def _poly_at(vals, lg: int, t: int) -> int:
    """value at t of the polynomial interpolating vals (ints, or a (d, 4) Montgomery
    array) on the size-2^lg domain"""
    arr = vals if isinstance(vals, np.ndarray) else O.ints_to_array(vals, R)
    coeffs = O.ntt(arr, lg, 1)  # ifft
    return O.unmont(O.poly_eval(coeffs, O.mont(t, R)), R)


def lagrange_at(k: int, lg: int, t: int) -> int:
    d = 1 << lg
    w = pow(O.ROOT28, 1 << (28 - lg), R)
    wk = pow(w, k, R)
    return (pow(t, d, R) - 1) * pow(d, -1, R) * wk * pow((t - wk) % R, -1, R) % R


def closed_form(inst: dict):
    """(scalars a, b, c, ic) and the expected affine proof points (packed u64 arrays)."""
    nc, ni, nw, lg = (int(x) for x in inst["header"][:4])
    td = inst["trapdoor"]
    t, al, be, ga, de = td["tau"], td["alpha"], td["beta"], td["gamma"], td["delta"]
    r = O.unmont(O.from_words(inst["r"]), R)
    s = O.unmont(O.from_words(inst["s"]), R)
    z = [O.unmont(O.from_words(v), R) for v in inst["z"][:ni]]
    ea, eb, ec = eval_vectors_mont(inst)
    At, Bt, Ct = (_poly_at(v, lg, t) for v in (ea, eb, ec))
    # u_i, v_i, w_i at t for the instance variables (their columns and input rows)
    u = [0] * ni
    v = [0] * ni
    w = [0] * ni
    for m, acc in (("a", u), ("b", v), ("c", w)):
        off, col = inst[f"{m}_off"], inst[f"{m}_col"]
        rows = np.repeat(np.arange(nc), np.diff(off.astype(np.int64)))
        for e in np.nonzero(col < ni)[0]:
            cf = O.unmont(O.from_words(inst[f"{m}_coef"][e]), R)
            acc[int(col[e])] = (acc[int(col[e])] + cf * lagrange_at(int(rows[e]), lg, t)) % R
    for i in range(ni):
        u[i] = (u[i] + lagrange_at(nc + i, lg, t)) % R
    ic = sum(z[i] * (be * u[i] + al * v[i] + w[i]) for i in range(ni)) % R
    a = (al + At + r * de) % R
    b = (be + Bt + s * de) % R
    c = ((be * At + al * Bt - ic + At * Bt) * pow(de, -1, R) + s * a + r * b - r * s * de) % R
    return {"a": a, "b": b, "c": c, "ic": ic, "alpha": al, "beta": be, "delta": de,
            "At": At, "Bt": Bt, "Ct": Ct}


def points_of(cf: dict):
    """Affine points a G1, b G2, c G1 as packed u64 arrays (8, 16, 8 words)."""
    g1 = O.g1_pack([O.G1_GEN])[0]
    g2 = O.g2_pack([O.G2_GEN])[0]
    A = np.zeros(8, np.uint64)
    B = np.zeros(16, np.uint64)
    C = np.zeros(8, np.uint64)
    O.lib().or_g1_mul_aff(O._ptr(A), O._ptr(g1), O._ptr(np.array(O.to_words(cf["a"]), np.uint64)))
    O.lib().or_g2_mul_aff(O._ptr(B), O._ptr(g2), O._ptr(np.array(O.to_words(cf["b"]), np.uint64)))
    O.lib().or_g1_mul_aff(O._ptr(C), O._ptr(g1), O._ptr(np.array(O.to_words(cf["c"]), np.uint64)))
    return A, B, C

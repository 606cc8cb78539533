"""The stencil-layout planner (gip_amd/csrc/gip_layout.h) on the CPU.

tests/layout_check.cpp plans the layout of a CSR operator and forms y = A x
from it exactly as the device kernels do (slots in order with clamped
columns, then the row's tail); the products must equal the CSR products in
stored order bitwise, every stored entry must land in exactly one value
slot, and every tail column must lie inside the operator.  The symmetric
form of the slot values (plan_symmetric, used by the PCG's product) must
pass its check on the generated Laplacians, fail it when one value is
perturbed, and give the full-layout slot sums bitwise whenever it passes.
"""
import json
import shutil
import subprocess

import numpy as np
import pytest

from gip_amd import cavity

from .conftest import ROOT


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("layout")
    exe = d / "layout_check"
    subprocess.run([gxx, "-O2", "-std=c++17", "-o", str(exe), str(ROOT / "tests" / "layout_check.cpp")], check=True)
    return exe, d


def _run(checker, rowptr, col, val, seed=0, frac=0.05):
    exe, d = checker
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    ncx = int(col.max()) + 1 if col.size else 1
    x = np.random.default_rng(seed).standard_normal(ncx)
    for name, a in (("rp", rowptr), ("col", col), ("val", val), ("x", x)):
        a.tofile(d / name)
    out = subprocess.run([str(exe), str(rowptr.size - 1), str(val.size), str(d / "rp"), str(d / "col"),
                          str(d / "val"), str(d / "x"), str(ncx), str(frac)], capture_output=True, text=True,
                         check=True)
    return json.loads(out.stdout)


def _clean(r):
    return r["bad_pos"] == 0 and r["bad_cols"] == 0 and r["bad_rows"] == 0


@pytest.mark.parametrize("n,beta", [(24, 0.0), (24, 1.5), (46, 0.0), (70, 0.0)])
def test_cavity_operators_plan_bitwise(checker, n, beta):
    c = cavity.generate(n, beta=beta, gamma=1e-2)
    planned = {}
    for name in cavity.OPS:
        A = c.op(name)
        r = _run(checker, A.rowptr, A.col, A.val)
        planned[name] = r["planned"]
        if r["planned"]:
            assert _clean(r), (name, r)
            assert r["w"] <= 8
    # the 7-point operators always fit, with a tail of boundary entries only
    for name in ("L", "D"):
        assert planned[name], name
    r = _run(checker, c.op("L").rowptr, c.op("L").col, c.op("L").val)
    ghosts = c.n_all - c.n_inner
    assert r["tail"] <= ghosts + 64 * 8, r


def test_laplacian_streams_fewer_bytes_than_ell(checker):
    """The byte count the automatic choice compares (values, tail columns,
    tail words, slice headers) against 12 bytes per ELL entry: the 7-point
    Laplacian (7 entries in every row) needs about two thirds."""
    c = cavity.generate(64, gamma=1e-3, ops=cavity.OP_BITS["L"])
    L = c.op("L")
    assert np.all(np.diff(L.rowptr) == 7)
    r = _run(checker, L.rowptr, L.col, L.val)
    assert r["planned"] and _clean(r) and r["w"] == 7, r
    ell = 12 * 7 * L.nrows
    assert r["bytes"] < 0.72 * ell, (r["bytes"], ell)


def _banded(n, offsets, seed, extra=0.02, empty=(), shuffle_rows=0):
    rng = np.random.default_rng(seed)
    rowptr, cols, vals = [0], [], []
    shuffled = set(rng.choice(n, size=shuffle_rows, replace=False).tolist()) if shuffle_rows else set()
    for r in range(n):
        c = [r + d for d in offsets if 0 <= r + d < n]
        if r in empty:
            c = []
        elif rng.random() < extra:
            c.insert(int(rng.integers(0, len(c) + 1)), int(rng.integers(0, n + 100)))
        if r in shuffled:
            rng.shuffle(c)
        cols += c
        vals += list(rng.standard_normal(len(c)))
        rowptr.append(len(cols))
    return np.array(rowptr), np.array(cols), np.array(vals)


@pytest.mark.parametrize("seed", range(4))
def test_tails_empty_rows_and_unsorted_rows(checker, seed):
    rp, cl, vl = _banded(5000, (-70, -1, 0, 1, 70), seed, empty={0, 64, 65, 4999}, shuffle_rows=20)
    r = _run(checker, rp, cl, vl, seed)
    assert r["planned"] and _clean(r), r
    assert r["tail"] > 0


def test_duplicate_entries_keep_their_order(checker):
    # rows with the diagonal stored twice, around another entry
    n = 512
    rp, cl = [0], []
    dup_rows = 0
    for i in range(n):
        row = [i, min(i + 1, n - 1), i] if i % 3 == 0 else [max(i - 1, 0), i, min(i + 1, n - 1)]
        dup_rows += len(set(row)) < len(row)
        cl += row
        rp.append(len(cl))
    vl = np.random.default_rng(3).standard_normal(len(cl))
    # a third of the rows put their second diagonal entry in the tail: allow it
    r = _run(checker, rp, cl, vl, frac=0.5)
    assert r["planned"] and _clean(r), r
    assert r["tail"] == dup_rows  # each repeated entry, and only it, goes to the tail


def test_irregular_operator_is_not_planned(checker):
    rng = np.random.default_rng(7)
    n, per = 4000, 6
    rp = np.arange(0, n * per + 1, per)
    cl = np.sort(rng.integers(0, n, size=(n, per)), axis=1).ravel()
    r = _run(checker, rp, cl, rng.standard_normal(n * per))
    assert not r["planned"]


@pytest.mark.parametrize("n,beta", [(24, 0.0), (24, 1.5), (46, 0.0)])
def test_symmetric_form_of_the_laplacian(checker, n, beta):
    c = cavity.generate(n, beta=beta, gamma=1e-2)
    L = c.op("L")
    r = _run(checker, L.rowptr, L.col, L.val)
    assert r["planned"] and _clean(r), r
    # paired offsets 0, 1, n, n^2, and arrays for the boundary slices' ghost slots
    assert r["sym_planned"] and r["sym_nd"] == 4 and 4 <= r["sym_nu"] <= 8, r
    assert r["sym_ok"] and r["sym_bad_rows"] == 0, r


def test_symmetric_form_refuses_a_perturbed_value(checker):
    c = cavity.generate(24, gamma=1e-2)
    L = c.op("L")
    val = L.val.copy()
    # one off-diagonal entry of an interior row, changed in its last bit
    r0 = 24 * 24 * 12 + 24 * 12 + 12
    js = [j for j in range(L.rowptr[r0], L.rowptr[r0 + 1]) if L.col[j] == r0 + 24]
    assert js
    val[js[0]] = np.nextafter(val[js[0]], 0.0)
    r = _run(checker, L.rowptr, L.col, val)
    assert r["planned"] and _clean(r) and r["sym_planned"], r
    assert not r["sym_ok"], r


def test_symmetric_form_passes_only_where_it_is_exact(checker):
    """Whatever the operator, a passed check means bitwise equal slot sums."""
    c = cavity.generate(16, beta=1.5, gamma=1e-2)
    for name in cavity.OPS:
        A = c.op(name)
        r = _run(checker, A.rowptr, A.col, A.val)
        if r["planned"] and r["sym_planned"] and r["sym_ok"]:
            assert r["sym_bad_rows"] == 0, (name, r)


def test_one_entry_tail_records_at_128(checker):
    """From 128 cells per line on, a 64-row slice holds at most one
    x-boundary row: its one ghost entry is the slice's whole tail, and the
    compact record carries its lane, position and column (kind 3); the sums
    stay bitwise."""
    c = cavity.generate(128, gamma=1e-3, ops=cavity.OP_BITS["L"])
    L = c.op("L")
    r = _run(checker, L.rowptr, L.col, L.val)
    assert r["planned"] and _clean(r), r
    # the first and last slice of every line off the y and z boundaries
    # (those lines' slices also hold their y/z ghosts' entries)
    assert r["one_entry_slices"] == 2 * 126 * 126, r

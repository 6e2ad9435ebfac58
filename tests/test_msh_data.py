"""gmsh `$ElementData` input and MSH output: the library's reader and writer
(fl_element_data_*, fl_msh_write_*) against the reference's own unit-test
expectations (test_units/fields/field_elementwise_test.cpp with
simplest_cube_data.msh), the reference's data files (tests/18 init file,
tests/01 and tests/13 output files, byte for byte) and, on generated files,
against the oracle's restatement of the reading rules (oracle/msh_data.py).
Host code only: no GPU is needed."""
import math
import os
import random

import numpy as np
import pytest

import cases
import flowgpu as G
import msh_data as OM

CUBE = cases.ref_path("simplest_cube.msh")
CUBE_DATA = cases.ref_path("simplest_cube_data.msh")


def _oracle_ids(mesh, boundary):
    return [int(i) for i in G.element_ids(mesh, boundary)]


# field_elementwise_test.cpp:95-198: (field, n_comp, boundary, expected rows, rows that stay zero)
UNIT_CASES = [
    ("scalar", 1, False, [[(i + 1) * 0.1] for i in range(9)], 0),
    ("scalar", 1, True, [[1.0 + (i + 1) * 0.1] for i in range(4)], 14),
    ("vector_fixed", 3, False, [[1, 2, 3]] * 9, 0),
    ("vector_fixed", 3, True, [[4, 5, 6]] * 4, 14),
    ("tensor_fixed", 4, False, [[1, 2, 3, 4]] * 9, 0),
    ("tensor_fixed", 4, True, [[4, 5, 6, 7]] * 4, 14),
]


@pytest.mark.parametrize("field,nc,bnd,rows,nzero", UNIT_CASES)
def test_reference_unit_expectations(field, nc, bnd, rows, nzero):
    m = G.Mesh(CUBE)
    f = G.FieldElementwise(m, CUBE_DATA, field, n_comp=nc, boundary=bnd)
    assert f.set_time(0.0)
    want = np.array(rows + [[0.0] * nc] * nzero, dtype=np.float64).reshape(-1)
    got = f.value().reshape(-1)
    # EXPECT_DOUBLE_EQ: within 4 units in the last place ((i+1)*0.1 against the parsed "0.3")
    assert np.all(np.abs(got - want) <= 4 * np.spacing(np.abs(want)))
    # the oracle restatement gives the same on the same file
    r = OM.ElementDataReader(CUBE_DATA)
    d = [0.0] * (len(want))
    assert r.read(field, 0.0, nc, _oracle_ids(m, bnd), d)
    assert np.array_equal(np.array(d), got)


def test_bulk_ids_end_section_with_warning():
    """Rows 10..13 of simplest_cube_data.msh are boundary elements: the bulk read
    stops at the first of them (msh_gmshreader.cc:339-343) and the entity count
    differs from the file's (:384-388)."""
    m = G.Mesh(CUBE)
    f = G.FieldElementwise(m, CUBE_DATA, "scalar")
    f.set_time(0.0)
    n_read, warns = f.info()
    assert n_read == 9
    assert any("Data ID 10 not found" in w for w in warns)
    assert any("Wrong number of entities: 13, using 9" in w for w in warns)


def test_time_frames_and_tolerance(tmp_path):
    """Sections apply in file order while time <= t (1 + 2 eps); the last one
    stays for all later times; +inf changes nothing."""
    m = G.Mesh(CUBE)
    ids = _oracle_ids(m, False)
    p = tmp_path / "frames.msh"
    with open(p, "w") as fh:
        for t, base in ((0.0, 10.0), (1.0, 20.0), (2.0, 30.0)):
            fh.write('$ElementData\n1\n"k"\n1\n%r\n3\n0\n1\n%d\n' % (t, len(ids)))
            for i in ids:
                fh.write("%d %r\n" % (i, base + i))
            fh.write("$EndElementData\n")
    f = G.FieldElementwise(m, str(p), "k")
    r = OM.ElementDataReader(str(p))
    d = [0.0] * len(ids)
    below_one = float(np.nextafter(1.0, 0.0))
    seq = [(0.0, True, 10.0), (0.5, False, 10.0), (below_one, True, 20.0), (1.0, False, 20.0),
           (1.5, False, 20.0), (7.0, True, 30.0), (9.0, False, 30.0), (math.inf, False, 30.0)]
    for t, upd, base in seq:
        assert f.set_time(t) == upd, t
        if not math.isinf(t):
            assert r.read("k", t, 1, ids, d) == upd, t
            assert np.array_equal(np.array(d), f.value())
        assert np.array_equal(f.value(), base + np.array(ids, np.float64)), t


def test_sparse_ids_test18_init():
    """tests/18_diffusion_fracture/input/test18_init.msh on test18.msh: element
    ids 1, 9, 17, 25..88; concentration 1 on element 45 only."""
    m = G.Mesh(cases.ref_path("test18.msh"))
    f = G.FieldElementwise(m, cases.ref_path("test18_init.msh"), "concentration")
    assert f.set_time(0.0)
    v = f.value()
    ids = G.element_ids(m)
    assert list(ids[:4]) == [1, 9, 17, 25] and len(ids) == 67
    want = np.where(ids == 45, 1.0, 0.0)
    assert np.array_equal(v, want)
    assert f.info()[0] == 67


def test_errors(tmp_path):
    m = G.Mesh(CUBE)
    with pytest.raises(G.FlowError, match="Can not open"):
        G.FieldElementwise(m, str(tmp_path / "none.msh"), "scalar")
    f = G.FieldElementwise(m, CUBE_DATA, "no_such_field")
    with pytest.raises(G.FlowError, match="missing '\\$ElementData' section"):
        f.set_time(0.0)
    p = tmp_path / "bad.msh"
    p.write_text('$ElementData\n1\n"k"\n0\n3\n0\n1\n2\n1 0.5\n2 x\n$EndElementData\n')
    f = G.FieldElementwise(m, str(p), "k")
    with pytest.raises(G.FlowError, match="Wrong format of \\$ElementData line"):
        f.set_time(0.0)
    p.write_text('$ElementData\n1\n"k"\n0\n3\n0\n2\n1\n1 0.5\n$EndElementData\n')
    f = G.FieldElementwise(m, str(p), "k", n_comp=2)
    with pytest.raises(G.FlowError, match="Missing token"):
        f.set_time(0.0)
    # bulk ids that do not increase (mesh.cc:790)
    q = tmp_path / "dec.msh"
    q.write_text("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n3\n1 0 0 0\n2 1 0 0\n3 0 1 0\n$EndNodes\n"
                 "$Elements\n2\n5 1 2 1 1 1 2\n4 1 2 1 1 2 3\n$EndElements\n")
    md = G.Mesh(str(q))
    with pytest.raises(G.FlowError, match="non-increasing"):
        G.FieldElementwise(md, str(p), "k")


def _random_file(rng, path, ids_bulk, fields):
    """A file mixing sections of several fields at increasing times, with ids
    missing from the mesh, unknown ids, out-of-order rows, comment and blank
    lines, tabs, surplus tags and component mismatches."""
    lines = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat"]
    t = 0.0
    for _ in range(rng.randint(3, 9)):
        name = rng.choice(sorted(fields))
        t += rng.choice([0.0, 0.25, 0.5, 1.0])
        ncomp = fields[name] if rng.random() < 0.85 else rng.choice([1, 2, 3])
        pool = sorted(set(ids_bulk) | {max(ids_bulk) + k for k in range(1, 4)} | {0})
        rows = sorted(rng.sample(pool, rng.randint(0, len(pool))))
        if rows and rng.random() < 0.3:
            k = rng.randrange(len(rows))
            rows.insert(k, rows[-1] + rng.randint(0, 3))  # an id beyond the mesh ends the section
        n_ent = len(rows) if rng.random() < 0.8 else len(rows) + rng.randint(-1, 1)
        n_ent = max(0, min(n_ent, len(rows)))
        lines.append("$ElementData")
        st = ['"%s"' % name] + (['"interp"'] if rng.random() < 0.3 else []) + (['"x y"'] if rng.random() < 0.2 else [])
        lines += [str(len(st))] + st
        rt = [repr(t)] + (["7.5"] if rng.random() < 0.2 else [])
        lines += [str(len(rt))] + rt
        it = [str(rng.randint(0, 9)), str(ncomp), str(n_ent)] + (["0"] if rng.random() < 0.2 else [])
        lines += [str(len(it))] + it
        if rng.random() < 0.3:
            lines.append("# a comment")
        for r in rows[:n_ent]:
            vals = ["%.17g" % rng.uniform(-5, 5) for _ in range(ncomp)]
            sep = "\t " if rng.random() < 0.2 else " "
            lines.append(sep.join([str(r)] + vals))
            if rng.random() < 0.05:
                lines.append("")
        lines.append("$EndElementData")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


@pytest.mark.parametrize("seed", range(40))
def test_reader_matches_oracle_generated(tmp_path, seed):
    rng = random.Random(seed)
    m = G.Mesh(CUBE)
    bulk = _oracle_ids(m, False)
    p = str(tmp_path / "g.msh")
    comps = {"a": 1, "b": 3}
    _random_file(rng, p, bulk, comps)
    for field in ("a", "b"):
        for bnd in (False, True):
            ids = _oracle_ids(m, bnd)
            nc = comps[field]
            # section times are multiples of 0.25: query on them, just below them and between
            grid = [0.25 * rng.randint(0, 24) for _ in range(rng.randint(1, 5))]
            times = sorted(rng.choice([g, float(np.nextafter(g, -1.0)), g + rng.uniform(0.0, 0.25)]) for g in grid)
            r = OM.ElementDataReader(p)
            d = [0.0] * (len(ids) * nc)
            f = G.FieldElementwise(m, p, field, n_comp=nc, boundary=bnd)
            for t in times:
                try:
                    want_upd = r.read(field, t, nc, ids, d)
                    want_err = None
                except ValueError as e:
                    want_err = e
                if want_err is not None:
                    with pytest.raises(G.FlowError):
                        f.set_time(t)
                    break
                assert f.set_time(t) == want_upd
                assert np.array_equal(f.value().reshape(-1), np.array(d)), (field, bnd, t)
                n_read, warns = f.info()
                if want_upd:
                    assert n_read == r.n_read
                assert sum("not found or is not in order" in w for w in warns) == r.warnings.count("skip-rest")
                assert sum("number of components" in w for w in warns) == r.warnings.count("components")
                assert sum("number of entities" in w for w in warns) == r.warnings.count("entities")


# ---- writer -----------------------------------------------------------------------


def _sections(text):
    out = []
    for s in text.split("$ElementData\n")[1:]:
        L = s.split("\n")
        name, units = L[1].strip('"')[:-1].split("_[")
        nc, n = int(L[6]), int(L[7])
        v = np.array([[float(x) for x in l.split()[1:]] for l in L[8:8 + n]])
        out.append((name, units, float(L[3]), int(L[5]), v[:, 0] if nc == 1 else v))
    return out


def test_writer_reproduces_test13_transport_output(tmp_path):
    """tests/13 transport.msh: mesh head and 201 concentration sections, written
    from the file's own values, equal the reference's file byte for byte."""
    gold = open(cases.ref_path("test13_transport.msh")).read()
    m = G.Mesh(cases.ref_path("test13.msh"))
    out = G.OutputMSH(m, str(tmp_path / "t.msh"))
    secs = _sections(gold)
    assert len(secs) == 201
    for name, units, t, step, v in secs:
        assert step == out.current_step
        out.write_data(t, [(name, units, v)])
    assert open(tmp_path / "t.msh").read() == gold


def test_writer_reproduces_test01_output(tmp_path):
    """tests/01 flow_gmsh.con/test1.msh: head plus pressure, piezometric head and
    3-component velocity sections (rows end with a blank) byte for byte; the
    file's $NodeData section is not produced by this library."""
    import re
    gold = open(cases.ref_path("test01_output.msh")).read()
    gold_el = re.sub(r"\$NodeData.*?\$EndNodeData\n", "", gold, flags=re.S)
    m = G.Mesh(cases.ref_path("test01_mesh.msh"))
    out = G.OutputMSH(m, str(tmp_path / "t.msh"))
    out.write_head()
    for name, units, t, step, v in _sections(gold):
        out.write_elem_data(name, units, v, t, step)
    assert open(tmp_path / "t.msh").read() == gold_el


def test_written_data_reads_back_exactly(tmp_path):
    """Writer and reader close the loop: 17 significant digits round-trip every
    double, bulk and 3-component data, at two output times, on a mesh whose
    element ids are 1..E (test13.msh)."""
    m = G.Mesh(cases.ref_path("test13.msh"))
    E = m.n_elements
    rng = np.random.default_rng(5)
    a0, a1 = rng.standard_normal(E) * 1e3, rng.standard_normal(E) ** 3
    v0 = rng.standard_normal((E, 3)) * 1e-7
    p = str(tmp_path / "o.msh")
    out = G.OutputMSH(m, p, precision=17)
    out.write_data(0.0, [("k", "m/s", a0), ("v", "L/T", v0)])
    out.write_data(2.5, [("k", "m/s", a1), ("v", "L/T", v0 * 2)])
    f = G.FieldElementwise(m, p, "k_[m/s]")
    assert f.set_time(1.0)
    assert np.array_equal(f.value(), a0)
    assert f.set_time(2.5)
    assert np.array_equal(f.value(), a1)
    g = G.FieldElementwise(m, p, "v_[L/T]", n_comp=3)
    assert g.set_time(10.0)
    assert np.array_equal(g.value(), v0 * 2)
    # the default precision is the reference's %g
    out6 = G.OutputMSH(m, str(tmp_path / "o6.msh"))
    out6.write_data(0.0, [("k", "m/s", a0)])
    rows = open(tmp_path / "o6.msh").read().split("$ElementData\n")[1].split("\n")[8:8 + E]
    assert rows == ["%d %g" % (i + 1, x) for i, x in enumerate(a0)]


def test_output_rows_are_element_indices(tmp_path):
    """Output rows carry index + 1, not the element's gmsh id (output_msh.cc:103,
    128), so reading an output back on a mesh with other ids matches rows by id:
    element j gets row ids[j] when ids[j] <= E and keeps zero otherwise."""
    m = G.Mesh(cases.ref_path("test01_mesh.msh"))
    E = m.n_elements
    ids = G.element_ids(m)
    a = np.arange(1.0, E + 1.0) * 0.5
    p = str(tmp_path / "o.msh")
    G.OutputMSH(m, p).write_data(0.0, [("k", "-", a)])
    f = G.FieldElementwise(m, p, "k_[-]")
    f.set_time(0.0)
    want = np.where(ids <= E, 0.5 * ids, 0.0)
    assert not np.array_equal(ids, np.arange(1, E + 1))
    assert np.array_equal(f.value(), want)


def test_writer_errors(tmp_path):
    m = G.Mesh(CUBE)
    with pytest.raises(G.FlowError):
        G.OutputMSH(m, str(tmp_path / "no_dir" / "x.msh")).write_head()
    out = G.OutputMSH(m, str(tmp_path / "p.msh"), precision=40)
    with pytest.raises(G.FlowError, match="precision"):
        out.write_elem_data("k", "-", np.zeros(9), 0.0, 0)

"""Ensemble ingestion: the native multi-model mmCIF / PDB reader against the
oracle's restatement of the reference's reading, and both against the
reference's own tests and fixtures.

Pins (reference file:line):
* counts: 2gb1 as mmCIF and as PDB 855 atoms / 56 residues / 1 chain / 1 model
  (bioshell-pdb/tests/test_cif_loading.rs:14-21, test_loading_pdb.rs:11-18),
  2fdo 1456 atoms / 2 chains / 214 residues (test_cif_loading.rs:24-31),
  2jqb 20 models of 255 atoms in 1 chain (test_structure.rs:156-166);
* one atom: 1akd fragment, residue A:515 atom " K  " has serial 3209
  (test_structure.rs:128-138);
* the 15 ATOM lines of test_pdb_atom_filters.rs:11-26 (3 residues, chains);
* PdbAtom::from_atom_line fields and its Display round trip
  (test_pdb_atom.rs:28-53);
* format_atom_name (pdb_atom.rs:294-305), is_hydrogen (pdb_atom.rs:338-350),
  quote_style (split_into_strings.rs:151-166), the ';' line iterator
  (cif_line_iterator.rs:45-86) and the CIF reader's edge cases
  (bioshell-cif/tests/test_cif.rs:110-200);
* the same structure read from 2gb1.cif and from 2gb1.pdb gives the same
  coordinates (two independent parsers of the reference's two formats).
"""
import os

import numpy as np
import pytest

import bioshell_amd as B
from bioshell_amd.crmsd import BioshellGpuError
from oracle import ingest_oracle as I
from tests.conftest import GOLDEN

SDIR = os.path.join(GOLDEN, "structures")
FIXTURES = sorted(f for f in os.listdir(SDIR) if f.endswith(".gz"))
SELECTIONS = ("all", "ca", "cb", "not_hydrogen")


def text_of(name):
    return I.read_text(os.path.join(SDIR, name))


def fields(atoms):
    return [(a.serial, a.name, a.alt_loc, a.res_name, a.chain_id, a.res_seq, a.i_code, a.occupancy,
             a.temp_factor, a.element, a.is_hetero_atom) for a in atoms]


def both(text, select="ca", chain=None, fmt="auto", n_threads=0):
    """Reads with the oracle and natively; asserts they agree (same ensemble,
    or both an error) and returns the native ensemble (None on error)."""
    try:
        want = I.ensemble(text, fmt=fmt, selection=select, chain=chain)
        werr = None
    except I.IngestError as e:
        want, werr = None, e
    try:
        got = B.Ensemble.from_text(text, select=select, chain=chain, format=fmt, n_threads=n_threads)
        gerr = None
    except BioshellGpuError as e:
        got, gerr = None, e
    if werr is not None or gerr is not None:
        assert werr is not None and gerr is not None, f"oracle: {werr!r}, native: {gerr!r}"
        assert gerr.status == -9, gerr
        return None
    assert got.xyz.shape == want.xyz.shape
    assert np.array_equal(got.xyz, want.xyz, equal_nan=True)
    assert fields(got.atoms) == fields(want.atoms)
    assert (got.n_structure_atoms, got.n_chains, got.n_residues) == \
        (want.n_structure_atoms, want.n_chains, want.n_residues)
    return got


# ------------------------------------------------------------ reference pins

@pytest.mark.parametrize("name", ["2gb1.cif.gz", "2gb1.pdb.gz"])
def test_reference_counts_2gb1(name):
    e = both(text_of(name), select="all")
    assert (e.n_structure_atoms, e.count_residues(), e.count_chains(), e.count_models()) == (855, 56, 1, 1)


def test_reference_counts_2fdo():
    e = both(text_of("2fdo.cif.gz"), select="all")
    assert (e.n_structure_atoms, e.count_chains(), e.count_models(), e.count_residues()) == (1456, 2, 1, 214)


def test_reference_counts_2jqb():
    e = both(text_of("2jqb.cif.gz"), select="all")
    assert e.count_models() == 20 and e.n_atoms == 255 and e.count_chains() == 1
    ca = both(text_of("2jqb.cif.gz"), select="ca")
    assert ca.xyz.shape == (20, 15, 3)  # 13 ATOM residues + 2 hydroxyprolines (HETATM HYP)


def test_reference_atom_1akd():
    e = both(text_of("1akd_fragm.pdb.gz"), select="all")
    hits = [a for a in e.atoms if a.chain_id == "A" and a.res_seq == 515 and a.name == " K  "]
    assert len(hits) == 1 and hits[0].serial == 3209


FILTER_LINES = [
    f"ATOM    {s}  {n} ALA {c}  {r}      {xyz}  1.00 {b}           {el}"
    for c, r in (("A", 68), ("A", 69), ("B", 68))
    for s, n, xyz, b, el in (
        (514, "N  ", "26.532  28.200  28.365", "17.85", "N"),
        (515, "CA ", "25.790  28.757  29.513", "16.12", "C"),
        (516, "C  ", "26.891  29.054  30.649", "15.28", "C"),
        (517, "O  ", "26.657  29.867  31.341", "20.90", "O"),
        (518, "CB ", "25.155  27.554  29.987", "21.91", "C"))]


def test_reference_filter_lines():
    text = "\n".join(FILTER_LINES) + "\n"
    assert both(text, select="all").count_residues() == 3
    assert both(text, select="ca").n_atoms == 3
    assert both(text, select="all", chain="A").n_atoms == 10
    assert both(text, select="all", chain="B").n_atoms == 5
    assert both(text, select="cb", chain="B").n_atoms == 1


PARSE_LINES = [
    "ATOM   2831  OE1BGLN A 294C    -27.117  12.343  28.479  1.00  9.58           O  ",
    "HETATM 3169  PB  BIS A 499     -10.950  43.372   1.324  1.00 18.00           P  ",
]


@pytest.mark.parametrize("line", PARSE_LINES)
def test_reference_atom_line_round_trip(line, tmp_path):
    e = both(line + "\n", select="all")
    a = e.atoms[0]
    if line.startswith("ATOM"):
        assert (a.serial, a.alt_loc, a.res_name, a.chain_id, a.res_seq, a.i_code) == (2831, "B", "GLN", "A", 294, "C")
        assert tuple(e.xyz[0, 0]) == (-27.117, 12.343, 28.479)
        assert (a.occupancy, a.temp_factor, a.element, a.is_hetero_atom) == (1.0, 9.58, "O", False)
    else:
        assert a.is_hetero_atom
    p = tmp_path / "one.pdb"
    e.write_pdb(str(p))
    assert p.read_text() == line + "\n"
    o = I.ensemble(line + "\n", selection="all")
    assert I.write_pdb(o.atoms, o.xyz[0]) == line + "\n"


def test_reference_format_atom_name():
    cases = [("CA", "C", " CA "), ("CD2", "C", " CD2"), ("CA", "CA", "CA  "), ("1HA", "H", "1HA "),
             ("OXT", None, " OXT"), ("H", "H", " H  ")]
    for name, el, want in cases:
        assert I.format_atom_name(name, el) == want
    # the native reader applies it to label_atom_id / type_symbol
    rows = [cif_row(k + 1, name, "ALA", 1, 0, 0, k, elem=el or "?") for k, (name, el, _) in enumerate(cases)]
    e = both(cif_text(rows), select="all")
    assert [a.name for a in e.atoms] == [w for _, _, w in cases]


def test_reference_predicates():
    a = I._pdb_atom("ATOM    320  CA  PHE A  43      16.101   9.057  19.587  1.00 18.18           C  ")
    assert not I.is_hydrogen(a)
    a.element = "H"
    assert I.is_hydrogen(a)
    a.element, a.name = None, "1HA"
    assert I.is_hydrogen(a)
    a.name = "CA"  # IsCA pads incorrectly padded names (pdb_atom_filters.rs:370-387)
    assert I.selected(a, "ca", None)
    # mercury (name "HG  ") counts as hydrogen, calcium (name "CA  ") is no alpha carbon
    rows = [cif_row(1, "HG", "HG", 1, 0, 0, 0, elem="HG", group="HETATM"),
            cif_row(2, "CA", "CA", 2, 1, 0, 0, elem="CA", group="HETATM"),
            cif_row(3, "CA", "GLY", 3, 2, 0, 0)]
    assert both(cif_text(rows), select="not_hydrogen").n_atoms == 2
    assert both(cif_text(rows), select="ca").n_atoms == 1


def test_reference_quote_style_and_line_iterator():
    for tok, style in [('"Hello"', "both"), ("'World", "begin"), ("Rust'", "end"), ("No", "none"),
                       ("'Smart'", "both"), ("'Different\"", "begin")]:
        assert I._quote_style(tok) == style
    for text in [";\nmultiline\n;\n", ";multiline\n;\n", "line1\nline2\nline3\n",
                 ";\nmultiline\n\nmultiline2\n;\n"]:
        joined = "".join(l + "\n" for l in I.cif_line_iterator(I.rust_lines(text)))
        assert joined.strip() == text.strip()


def test_reference_cif_reader_edge_cases():
    b = I.read_cif("data_loop\nloop_\n_symmetry_equiv_pos.site_id\n_symmetry_equiv_pos.as_xyz\n"
                   "1 -x,-y,-z\n2 x,y,z\n")
    assert len(b) == 1 and len(b[0].loops[0].rows) == 2
    b = I.read_cif("data_ALA\n_chem_comp.name\n;\nmultiline\n;\n    #")
    assert len(b) == 1 and len(b[0].items) == 1
    b = I.read_cif("data_1crr\n_struct.entry_id                  1CRR\n_struct.title\n"
                   "'THE SOLUTION STRUCTURE AND DYNAMICS OF RAS P21.'\n_struct.pdbx_model_details        ?\n")
    assert len(b) == 1 and b[0].items["_struct.title"].startswith("'THE SOLUTION")
    b = I.read_cif("data_2jnw\n_struct.entry_id                  2JNW\nloop_\n_atom_type.symbol\nC\nH\nN\n"
                   "_pdbx_nmr_refine.details\n;XPLOR-NIH was used.\n;\n")
    assert len(b) == 1 and len(b[0].loops) == 1 and "_pdbx_nmr_refine.details" in b[0].items
    # the same header patterns ahead of an _atom_site loop, natively
    head = ("data_2jnw\n_struct.entry_id                  2JNW\nloop_\n_atom_type.symbol\nC\nH\nN\n"
            "_pdbx_nmr_refine.details\n;XPLOR-NIH was used.\n;\n_struct.title\n'A B C'\n")
    both(cif_text([cif_row(1, "CA", "GLY", 1, 0, 0, 0)], header=head), select="all")


def test_cif_and_pdb_of_2gb1_agree():
    for sel in SELECTIONS:
        c = B.Ensemble.from_text(text_of("2gb1.cif.gz"), select=sel if sel != "not_hydrogen" else "all")
        p = B.Ensemble.from_text(text_of("2gb1.pdb.gz"), select=sel if sel != "not_hydrogen" else "all")
        assert np.array_equal(c.xyz, p.xyz)
        assert [a.res_seq for a in c.atoms] == [a.res_seq for a in p.atoms]


# ------------------------------------------------------------ native == oracle

@pytest.mark.parametrize("name", FIXTURES)
@pytest.mark.parametrize("select", SELECTIONS)
def test_fixture_native_matches_oracle(name, select):
    both(text_of(name), select=select)


@pytest.mark.parametrize("name,chain", [("6ins.cif.gz", "B"), ("2fdo.cif.gz", "B"), ("1c5n.cif.gz", "H"),
                                        ("5edw.cif.gz", "A"), ("2gb1.pdb.gz", "A"), ("2gb1.pdb.gz", "Z")])
def test_chain_filter(name, chain):
    e = both(text_of(name), select="ca", chain=chain)
    assert all(a.chain_id == chain for a in e.atoms)


def test_file_and_gz_reading(tmp_path):
    raw = text_of("2jqb.cif.gz")
    plain = tmp_path / "2jqb.cif"
    plain.write_bytes(raw)
    a = B.load_ensemble(str(plain), select="all")
    b = B.load_ensemble(os.path.join(SDIR, "2jqb.cif.gz"), select="all")
    assert np.array_equal(a.xyz, b.xyz) and fields(a.atoms) == fields(b.atoms)
    with pytest.raises(BioshellGpuError) as e:
        B.load_ensemble(str(tmp_path / "missing.cif"))
    assert e.value.status == -10


# ------------------------------------------------------------ edge cases

COLS = ["group_PDB", "id", "type_symbol", "label_atom_id", "label_alt_id", "label_comp_id", "label_asym_id",
        "label_entity_id", "label_seq_id", "pdbx_PDB_ins_code", "Cartn_x", "Cartn_y", "Cartn_z", "occupancy",
        "B_iso_or_equiv", "pdbx_formal_charge", "auth_seq_id", "auth_comp_id", "auth_asym_id",
        "auth_atom_id", "pdbx_PDB_model_num"]


def cif_row(serial, name, comp, seq, x, y, z, model=1, elem="C", alt=".", chain="A", ins="?", group="ATOM"):
    return [group, str(serial), elem, name, alt, comp, chain, "1", str(seq), ins, f"{x:.3f}", f"{y:.3f}",
            f"{z:.3f}", "1.00", "0.00", "?", str(seq), comp, chain, name, str(model)]


def cif_text(rows, header="data_test\n#\n", extra="", cols=COLS, sep=" ", eol="\n"):
    s = header + "loop_" + eol + "".join(f"_atom_site.{c}{eol}" for c in cols)
    s += "".join(sep.join(r) + eol for r in rows) + "#" + eol + extra
    return s


def models(n_models, n_atoms, numbers=None, drop=None):
    rng = np.random.default_rng(n_models * 100 + n_atoms)
    rows = []
    for m in range(n_models):
        num = numbers[m] if numbers else m + 1
        for k in range(n_atoms):
            if drop == (m, k):
                continue
            x, y, z = rng.uniform(-50, 50, 3)
            rows.append(cif_row(k + 1, ["N", "CA", "C", "O"][k % 4], "ALA", k // 4 + 1, x, y, z, model=num,
                                elem=["N", "C", "C", "O"][k % 4]))
    return rows


def test_multi_model_numbering_by_first_appearance():
    e = both(cif_text(models(3, 8, numbers=[5, 2, 9])), select="all")
    assert e.count_models() == 3 and e.n_atoms == 8


def test_model_atom_count_mismatch_is_an_error():
    assert both(cif_text(models(3, 8, drop=(2, 5))), select="all") is None


def test_interleaved_models_are_an_error():
    rows = models(2, 4)
    assert both(cif_text(rows[:2] + rows[4:6] + rows[2:4] + rows[6:]), select="all") is None


def test_crlf_and_tabs_and_unicode_whitespace():
    rows = models(2, 6)
    both(cif_text(rows, eol="\r\n"), select="all")
    both(cif_text(rows, sep="\t"), select="all")
    both(cif_text(rows, sep="   "), select="all")   # no-break space separates tokens
    both(cif_text(rows, sep=" "), select="all")      # em space


def test_text_field_with_atom_lines_before_the_loop():
    head = "data_x\n_struct.title\n;\nATOM 1 C CA . GLY A 1 1 ? 1 2 3 1 0 ? 1 GLY A CA 1\n;\n"
    e = both(cif_text(models(1, 4), header=head), select="all")
    assert e.n_atoms == 4


def test_rows_split_over_lines():
    rows = models(2, 6)
    text = "data_x\nloop_\n" + "".join(f"_atom_site.{c}\n" for c in COLS)
    for r in rows:
        text += " ".join(r[:7]) + "\n" + " ".join(r[7:]) + "\n"
    both(text, select="all")
    # one long line holding two rows
    text = "data_x\nloop_\n" + "".join(f"_atom_site.{c}\n" for c in COLS)
    text += " ".join(rows[0] + rows[1]) + "\n" + "".join(" ".join(r) + "\n" for r in rows[2:])
    both(text, select="all")


def test_text_field_value_inside_the_loop():
    rows = models(1, 3)
    text = "data_x\nloop_\n" + "".join(f"_atom_site.{c}\n" for c in COLS)
    r = rows[0]
    text += " ".join(r[:5]) + "\n;GLY\n;\n" + " ".join(r[6:]) + "\n"
    text += "".join(" ".join(r) + "\n" for r in rows[1:])
    both(text, select="all")


def test_quoted_tokens():
    rows = models(1, 4)
    rows[0][3] = "\"O5'\""
    rows[1][3] = "'N A'"
    rows[2][5] = "“X Y”"
    rows[3][3] = "'"
    rows[3].insert(4, "CB'")  # the lone quote absorbs the next word
    both(cif_text(rows), select="all")


def test_partial_last_row():
    cols = ["pdbx_PDB_model_num", "Cartn_x", "Cartn_y", "Cartn_z", "id", "label_atom_id", "label_alt_id",
            "label_comp_id", "auth_asym_id", "label_seq_id", "pdbx_PDB_ins_code", "occupancy",
            "B_iso_or_equiv", "type_symbol", "auth_seq_id", "label_entity_id", "extra1", "extra2"]
    full = ["1", "1.0", "2.0", "3.0", "1", "CA", ".", "GLY", "A", "1", "?", "1.0", "0.0", "C", "1", "1", "a", "b"]
    text = cif_text([full, full[:16]], cols=cols)  # the partial row holds every column the reader asks for
    e = both(text, select="all")
    assert e.n_atoms == 2
    assert both(cif_text([full, full[:15]], cols=cols), select="all") is None


def test_items_instead_of_a_loop():
    r = cif_row(7, "CA", "GLY", 3, 1.5, -2.25, 3.125)
    text = "data_x\n" + "".join(f"_atom_site.{c} {v}\n" for c, v in zip(COLS, r))
    e = both(text, select="all")
    assert e.n_atoms == 1 and tuple(e.xyz[0, 0]) == (1.5, -2.25, 3.125)


@pytest.mark.parametrize("mutate", ["coord", "model", "id", "alt", "occupancy", "missing_col", "two_blocks",
                                    "no_block", "dangling", "values_outside", "bad_item"])
def test_malformed_cif_is_an_error(mutate):
    rows = models(1, 3)
    cols = COLS
    head, extra = "data_test\n#\n", ""
    if mutate == "coord":
        rows[1][10] = "1.2.3"
    elif mutate == "model":
        rows[2][20] = "-1"
    elif mutate == "id":
        rows[0][1] = "x1"
    elif mutate == "alt":
        rows[0][4] = "AB"
    elif mutate == "occupancy":
        rows[1][13] = "1,0"
    elif mutate == "missing_col":
        cols = [c for c in COLS if c != "occupancy"]
        rows = [r[:13] + r[14:] for r in rows]
    elif mutate == "two_blocks":
        extra = "data_second\n_a.b c\n"
    elif mutate == "no_block":
        head = ""
    elif mutate == "dangling":
        extra = "_a.b\n_a.c\n"
    elif mutate == "values_outside":
        extra = "x y\n"
    elif mutate == "bad_item":
        extra = "_a.b c d\n"
    assert both(cif_text(rows, header=head, extra=extra, cols=cols), select="all") is None


def test_numbers_follow_rust_grammar():
    rows = models(1, 6)
    vals = ["+1.5", "-0.000", "1.", ".5", "1e3", "2.5E-2", "inf", "-Infinity", "NaN", "1e400",
            "123456789012345678901234567890", "0.1000000000000000055511151231257827"]
    for k, v in enumerate(vals):
        rows[k % 6][10 + k % 3] = v
    e = both(cif_text(rows), select="all")
    assert e is not None
    for bad in ["1_000", " 1", "e5", ".", "+", "1e", "0x10", "nan(1)"]:
        r = models(1, 2)
        r[0][10] = bad if " " not in bad else "'" + bad + "'"
        assert both(cif_text(r), select="all") is None, bad


def test_invalid_utf8():
    good = cif_text(models(2, 4)).encode()
    # an undecodable line after the loop: the mmCIF reader stops there
    both(good + b"# \xff\xfe\n", select="all")
    # inside the second model: the reading stops mid-model, which set_model rejects
    lines = good.split(b"\n")
    lines[-4] = lines[-4] + b" \xc3"
    both(b"\n".join(lines), select="all")
    pdb = ("\n".join(FILTER_LINES) + "\nREMARK \xe9\n").encode("latin-1")
    assert both(pdb, select="all") is None


def test_format_detection():
    assert both("", select="all") is None
    assert both("   \ndata_x\n", select="all") is None        # first non-empty line is blanks
    assert both("\n\nHEADER    x\n" + "\n".join(FILTER_LINES), select="all").n_atoms == 15
    assert both("# comment\ndata_x\n", select="all") is None


# ------------------------------------------------------------ PDB specifics

def pdb_models(n_models, n_atoms, with_model_records=True, drop=None):
    rng = np.random.default_rng(7 + n_models)
    out = []
    for m in range(n_models):
        if with_model_records:
            out.append(f"MODEL     {m + 1:4d}")
        for k in range(n_atoms):
            if drop == (m, k):
                continue
            x, y, z = rng.uniform(-99, 99, 3)
            rec = "HETATM" if k == n_atoms - 1 else "ATOM  "
            out.append(f"{rec}{k + 1:5d}  CA  ALA A{k + 1:4d}    {x:8.3f}{y:8.3f}{z:8.3f}  1.00  0.00           C  ")
        if with_model_records:
            out.append("ENDMDL")
    return "\n".join(out) + "\nEND\n"


def test_pdb_models():
    e = both(pdb_models(4, 7), select="ca")
    assert e.count_models() == 4 and e.n_atoms == 7 and e.atoms[-1].is_hetero_atom
    assert both(pdb_models(1, 5, with_model_records=False), select="all").count_models() == 1
    assert both(pdb_models(3, 5, drop=(1, 2)), select="all") is None
    # ENDMDL then coordinates without MODEL: the reference indexes past its models
    text = pdb_models(1, 3).replace("END\n", "") + "ATOM      1  CA  ALA A   1       0.000   0.000   0.000  1.00  0.00           C\n"
    assert both(text, select="all") is None
    # repeated MODEL records open one model
    text = "MODEL 1\nMODEL 1\n" + "\n".join(pdb_models(1, 3, with_model_records=False).split("\n")[:3]) + "\nENDMDL\n"
    assert both(text, select="all").count_models() == 1


def test_pdb_line_quirks():
    base = "ATOM      1  CA  ALA A   1      11.104  13.207   9.932  1.00 25.00           C"
    assert len(base) == 78 and base[76:78] == " C"
    assert both(base + "\n", select="all").atoms[0].element == "C"
    assert both(base[:77] + "\n", select="all").atoms[0].element is None  # 77 columns: no element
    # the element is read from column 78 on (pdb_atom.rs:98): iron "FE" reads as "E"
    assert both(base[:76] + "FE\n", select="all").atoms[0].element == "E"
    assert both(base[:60] + "\n", select="all") is None  # the first model needs 66 columns
    text = "MODEL 1\n" + base + "\nENDMDL\nMODEL 2\n" + base[:54] + "\nENDMDL\n"
    assert both(text, select="all").count_models() == 2  # later models need 54 columns
    assert both("MODEL 1\n" + base + "\nENDMDL\nMODEL 2\n" + base[:50] + "\nENDMDL\n", select="all") is None
    assert both(base.replace("  1.00", "  x.00") + "\n", select="all") is None


# ------------------------------------------------------------ threads, writer

def big_cif(n_models):
    raw = text_of("2jqb.cif.gz").decode()
    head, rest = raw.split("_atom_site.pdbx_PDB_model_num", 1)
    body, tail = rest.split("\n#", 1)
    lines = [l for l in body.split("\n") if l.strip()]
    first = [l for l in lines if l.split()[-1] == "1"]
    out = [head + "_atom_site.pdbx_PDB_model_num\n"]
    for m in range(n_models):
        for l in first:
            t = l.split()
            t[-1] = str(m + 1)
            t[10] = f"{float(t[10]) + m * 0.001:.3f}"
            out.append(" ".join(t) + "\n")
    return "".join(out) + "#" + tail


def test_threads_give_identical_results():
    text = big_cif(300).encode()   # ~6 MB: several chunks per thread
    one = B.Ensemble.from_text(text, select="all", n_threads=1)
    many = B.Ensemble.from_text(text, select="all", n_threads=8)
    assert one.count_models() == 300
    assert np.array_equal(one.xyz, many.xyz) and fields(one.atoms) == fields(many.atoms)
    want = I.ensemble(text, selection="ca")
    got = B.Ensemble.from_text(text, select="ca", n_threads=8)
    assert np.array_equal(got.xyz, want.xyz)


def test_writer_matches_oracle_and_reads_back(tmp_path):
    e = B.Ensemble.from_text(text_of("2jqb.cif.gz"), select="all")
    o = I.ensemble(text_of("2jqb.cif.gz"), selection="all")
    p = tmp_path / "2jqb.pdb"
    e.write_pdb(str(p))
    text = p.read_text()
    want = "".join(f"MODEL     {m + 1:4d}\n" + I.write_pdb(o.atoms, o.xyz[m]) + "ENDMDL\n" for m in range(20))
    assert text == want
    back = both(text, select="all")
    assert back.count_models() == 20 and np.array_equal(back.xyz, e.xyz)
    # threads on PDB input
    big = text * 15
    big = big.replace("ENDMDL\n", "ENDMDL\n")  # 300 models, same layout
    one = B.Ensemble.from_text(big, select="all", n_threads=1)
    many = B.Ensemble.from_text(big, select="all", n_threads=8)
    assert one.count_models() == 300 and np.array_equal(one.xyz, many.xyz)


def test_writer_renames_long_chain_ids(tmp_path):
    rows = [cif_row(1, "CA", "GLY", 1, 0, 0, 0, chain="AA"), cif_row(2, "CA", "GLY", 1, 1, 0, 0, chain="B"),
            cif_row(3, "CA", "GLY", 2, 2, 0, 0, chain="AA")]
    e = both(cif_text(rows), select="all")
    p = tmp_path / "r.pdb"
    e.write_pdb(str(p))
    o = I.ensemble(cif_text(rows), selection="all")
    assert p.read_text() == I.write_pdb(o.atoms, o.xyz[0])
    assert [l[21] for l in p.read_text().splitlines()] == ["A", "A", "B"]


def test_2gb1_pdb_written_back_verbatim(tmp_path):
    """Writing the parsed 2gb1.pdb reproduces its ATOM records byte for byte."""
    raw = text_of("2gb1.pdb.gz").decode()
    e = B.Ensemble.from_text(raw, select="all")
    p = tmp_path / "w.pdb"
    e.write_pdb(str(p))
    orig = [l for l in raw.splitlines() if l.startswith("ATOM")]
    got = p.read_text().splitlines()
    assert got == [l.ljust(80) for l in orig]

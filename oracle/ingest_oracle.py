"""CPU restatement of how BioShell reads a multi-model structure file into a
conformational ensemble.

TEST INFRASTRUCTURE ONLY: imported by ``tests/`` and by the ``cpu_baseline``
leg of ``bench.py`` as the checker of the native loader.  Nothing in the
product library imports it.

The reference reads an ensemble in two steps: ``Deposit::from_file`` builds a
``Structure`` whose atoms come from the first model and whose
``model_coordinates`` hold every model's positions; the caller then sets each
model in turn (``Structure::set_model``) and keeps the atoms that pass an atom
predicate (``IsCA`` and friends).  This module restates, in plain Python:

* mmCIF text -> loops: the line iterator that joins ``;`` text fields
  (bioshell-cif/src/cif_line_iterator.rs:16-40), the line classification
  (bioshell-cif/src/lib.rs:596-627), the block/loop state machine
  (lib.rs:653-760) with loop rows filled token by token (lib.rs:220-247,
  331-337), and the tokenizer that keeps quotes
  (bioshell-core/src/io/split_into_strings.rs:15-36, 70-125);
* ``_atom_site`` -> atoms and models: the first loop with a column starting
  ``_atom_site.`` (lib.rs:499-501), columns found by substring
  (bioshell-cif/src/column_mapping.rs:101-137, rows at 161-190), models
  numbered in order of first appearance with atoms taken while only one model
  has been seen (bioshell-pdb/src/pdb_atom.rs:130-170, fields at 175-203,
  names at 306-331);
* PDB text -> atoms and models: bioshell-pdb/src/load_pdb.rs:43-115 with
  ``PdbAtom::from_atom_line`` (pdb_atom.rs:87-116) and
  ``Vec3::from_pdb_line`` (bioshell-core/src/vec3.rs:241-246);
* format detection: load_cif.rs:155-167, load_pdb.rs:225-237, deposit.rs:81-85;
* counts and model selection: structure.rs:236-255, 277-285;
* predicates: pdb_atom_filters.rs:351-358 (name padding), 388-394 (IsCA),
  415-422 (IsCB), 191-199 (ByChain), pdb_atom.rs:352-365 (is_hydrogen);
* PDB writing: structure.rs:322-331, 828-850 and pdb_atom.rs:256-278.

Number parsing follows Rust's ``str::parse`` grammar for f64/i32/usize
(checked with regular expressions, then converted with Python's correctly
rounded ``float``/``int``).  Deposit metadata (header, entities, residue
types, helices, sheets) is not restated: it does not change the coordinates
or the atoms a predicate selects.

Pinned by tests/test_ingest.py against the reference's own tests and
fixtures (atom/residue/chain/model counts of 2gb1, 2fdo, 2jqb, one atom of
1akd, PdbAtom parsing and Display, format_atom_name, is_hydrogen,
quote_style, the ';' line iterator, the CIF reader's edge cases) and by
2gb1.cif and 2gb1.pdb giving identical coordinates.  Parity unpinned (no
reference test reaches them): Unicode whitespace and typographic quotes in
CIF tokens, loop rows spanning lines or holding ';' fields, _atom_site given
as items, the stop at a line that is not UTF-8.
"""
from __future__ import annotations

import gzip
import re
from dataclasses import dataclass

import numpy as np

# Rust's char::is_whitespace (White_Space property)
_RUST_WS = set(map(chr, [*range(0x09, 0x0E), 0x20, 0x85, 0xA0, 0x1680, *range(0x2000, 0x200B),
                           0x2028, 0x2029, 0x202F, 0x205F, 0x3000]))
_QUOTES = set("\"'“”‘’")

_F64 = re.compile(r"[+-]?(?:(?:[0-9]+\.?[0-9]*|\.[0-9]+)(?:[eE][+-]?[0-9]+)?|inf|infinity|nan)\Z",
                  re.IGNORECASE)
_I32 = re.compile(r"[+-]?[0-9]+\Z")
_USIZE = re.compile(r"\+?[0-9]+\Z")


class IngestError(ValueError):
    """The reference would return an error (or panic) on this input."""


def parse_f64(s: str) -> float:
    if not _F64.match(s):
        raise IngestError(f"invalid f64 {s!r}")
    return float(s)


def parse_i32(s: str) -> int:
    if not _I32.match(s):
        raise IngestError(f"invalid i32 {s!r}")
    v = int(s)
    if not -2 ** 31 <= v < 2 ** 31:
        raise IngestError(f"i32 out of range {s!r}")
    return v


def parse_usize(s: str) -> int:
    if not _USIZE.match(s):
        raise IngestError(f"invalid usize {s!r}")
    return int(s)


def rust_trim(s: str) -> str:
    a, b = 0, len(s)
    while a < b and s[a] in _RUST_WS:
        a += 1
    while b > a and s[b - 1] in _RUST_WS:
        b -= 1
    return s[a:b]


def rust_lines(data, stop_at_invalid: bool = False) -> list[str]:
    """BufRead::lines: split on '\\n'; a '\\r' just before the '\\n' goes too
    (an unterminated last line keeps its '\\r').  A line that is not UTF-8 is
    an error, or -- for the mmCIF reader, whose line iterator stops at the
    first error (cif_line_iterator.rs:19) -- the end of the input."""
    if isinstance(data, str):
        data = data.encode()
    if data == b"":
        return []
    parts = data.split(b"\n")
    last = parts.pop()
    raw = [p[:-1] if p.endswith(b"\r") else p for p in parts]
    if last != b"":
        raw.append(last)
    out = []
    for r in raw:
        try:
            out.append(r.decode("utf-8"))
        except UnicodeDecodeError:
            if stop_at_invalid:
                break
            raise IngestError("stream did not contain valid UTF-8") from None
    return out


# ------------------------------------------------------------------ CIF text

def cif_line_iterator(lines):
    """cif_line_iterator.rs:16-40: a line starting with ';' opens a text field
    that runs to the next line starting with ';'; the field is one item."""
    current = None
    for line in lines:
        if line.startswith(";"):
            if current is not None:
                yield current + line
                current = None
            else:
                current = line + "\n"
        elif current is not None:
            current += line + "\n"
        else:
            yield line
    if current is not None:
        yield current


def split_preserving_whitespace(s: str) -> list[str]:
    """split_into_strings.rs:15-36"""
    out = []
    start = 0
    in_ws = bool(s) and s[0] in _RUST_WS
    for i, c in enumerate(s):
        if (c in _RUST_WS) != in_ws:
            out.append(s[start:i])
            start = i
            in_ws = not in_ws
    out.append(s[start:])
    return out


def _quote_style(tok: str):
    if all(c in _RUST_WS for c in tok):
        return "ws"
    if len(tok.encode()) == 1 and tok in _QUOTES:  # str::len counts bytes
        return "begin"
    f, l = tok[0], tok[-1]
    if f == l and f in _QUOTES:
        return "both"
    if f in _QUOTES:
        return "begin"
    if l in _QUOTES:
        return "end"
    return "none"


def split_into_strings(s: str) -> list[str]:
    """split_into_strings.rs:70-95 with if_remove_quotes = false."""
    it = iter(split_preserving_whitespace(s))
    toks = [next(it)]
    last_space = ""
    for word in it:
        st = _quote_style(toks[-1])
        if st in ("none", "end", "both"):
            if not all(c in _RUST_WS for c in word):
                toks.append(word)
        elif st == "begin":
            toks[-1] = toks[-1] + last_space + word
        else:
            last_space = toks[-1]
    return toks


def classify(line: str):
    """CifLine::from_str, lib.rs:596-627."""
    t = rust_trim(line)
    if len(t) == 0 or t.startswith("#"):
        return ("empty", None)
    if t.startswith("data_"):
        return ("block", t[5:])
    if t.startswith("_"):
        item = split_into_strings(t)
        if len(item) == 1:
            return ("name", item[0])
        if len(item) == 2:
            return ("item", (item[0], item[1]))
        raise IngestError(f"invalid CIF line {line!r}")
    if t.startswith("loop_"):
        return ("loop", None)
    if t.startswith(";"):
        return ("text", t)
    return ("values", split_into_strings(t))


@dataclass
class CifLoop:
    columns: list
    rows: list

    def last_row_complete(self):
        return not self.rows or len(self.rows[-1]) == len(self.columns)

    def add_data(self, v):
        if self.last_row_complete():
            self.rows.append([v])
        else:
            self.rows[-1].append(v)

    def add_data_row(self, row):
        if self.last_row_complete() and len(row) == len(self.columns):
            self.rows.append(list(row))
        else:
            for v in row:
                self.add_data(v)


@dataclass
class CifBlock:
    name: str
    items: dict
    loops: list


def read_cif(text) -> list[CifBlock]:
    """read_cif_buffer, lib.rs:653-760."""
    blocks: list[CifBlock] = []
    loop = None
    open_item = None

    def close_loop():
        if not blocks:
            raise IngestError("loop outside a data block")
        blocks[-1].loops.append(loop)

    for line in cif_line_iterator(rust_lines(text, stop_at_invalid=True)):
        kind, val = classify(line)
        if kind == "loop":
            if loop is not None:
                close_loop()
            loop = CifLoop([], [])
        elif kind == "block":
            if loop is not None:
                close_loop()
                loop = None
            blocks.append(CifBlock(val, {}, []))
        elif kind == "item":
            if loop is not None:
                close_loop()
                loop = None
            if not blocks:
                raise IngestError("data item outside a data block")
            blocks[-1].items[val[0]] = val[1]
        elif kind == "text":
            if loop is not None:
                v = val[1:] if val.startswith(";") else val
                loop.add_data(v.rstrip(";"))
            elif open_item is not None:
                if not blocks:
                    raise IngestError("data item outside a data block")
                blocks[-1].items[open_item] = val
                open_item = None
            else:
                raise IngestError("multi-line string outside a data item")
        elif kind == "name":
            if loop is not None:
                if not loop.rows:
                    loop.columns.append(val)
                else:
                    close_loop()
                    loop = None
            if loop is None:
                if open_item is None:
                    open_item = val
                else:
                    raise IngestError(f"dangling data item {val}")
        elif kind == "values":
            if loop is not None:
                loop.add_data_row(val)
            elif open_item is not None and len(val) == 1:
                if not blocks:
                    raise IngestError("data item outside a data block")
                blocks[-1].items[open_item] = val[0]
                open_item = None
            else:
                raise IngestError(f"data values outside a loop: {line!r}")
    if loop is not None:
        close_loop()
    return blocks


ATOM_SITE_COLUMNS = ("id", "label_atom_id", "label_alt_id", "label_comp_id", "auth_asym_id",
                     "label_seq_id", "pdbx_PDB_ins_code", "Cartn_x", "Cartn_y", "Cartn_z",
                     "occupancy", "B_iso_or_equiv", "type_symbol", "pdbx_PDB_model_num",
                     "auth_seq_id", "label_entity_id")


def atom_site_rows(block: CifBlock):
    """CifTable::new + iter (column_mapping.rs:101-137, 161-190)."""
    prefix = "_atom_site."
    loop = next((l for l in block.loops if any(c.startswith(prefix) for c in l.columns)), None)
    if loop is not None:
        idx = []
        for col in ATOM_SITE_COLUMNS:
            pos = next((k for k, c in enumerate(loop.columns) if col in c), None)
            if pos is None:
                raise IngestError(f"missing CIF loop key {col}")
            idx.append(pos)
        for row in loop.rows:
            if any(k >= len(row) for k in idx):
                raise IngestError("incomplete row in _atom_site loop")
            yield [row[k] for k in idx]
        return
    vals = []
    for col in ATOM_SITE_COLUMNS:
        key = prefix + col
        if key not in block.items:
            raise IngestError(f"missing CIF loop key {key}")
        vals.append(block.items[key])
    yield vals


# ------------------------------------------------------------------ atoms

@dataclass
class Atom:
    serial: int
    name: str
    alt_loc: str
    res_name: str
    chain_id: str
    res_seq: int
    i_code: str
    occupancy: float
    temp_factor: float
    element: str | None
    is_hetero_atom: bool


def format_atom_name(atom_name: str, element: str | None) -> str:
    """pdb_atom.rs:306-331"""
    atom_name = rust_trim(atom_name)
    element = rust_trim(element or "")
    if atom_name == "":
        raise IngestError("empty atom name")
    res = [" "] * 4
    if atom_name[0].isdigit() and atom_name[0].isascii():
        for i, c in enumerate(atom_name[:4]):
            res[i] = c
    else:
        elen = 1 if element == "" else min(len(element.encode()), 2)
        pad = 1 if elen == 1 else 0
        for i, c in enumerate(atom_name[:4 - pad]):
            res[pad + i] = c
    return "".join(res)


def _single_char(tok: str, default: str) -> str:
    """value_or_default::<char> (bioshell-cif/src/lib.rs:799-801)."""
    if tok in ("?", "."):
        return default
    if len(tok) != 1:
        raise IngestError(f"not a single character: {tok!r}")
    return tok


def _cif_atom(t) -> Atom:
    """create_pdb_atom, pdb_atom.rs:175-203"""
    element = t[12]
    res_seq = parse_i32(t[14]) if t[14] not in ("?", ".") else parse_i32(t[5])
    return Atom(serial=parse_i32(t[0]), name=format_atom_name(t[1], element),
                alt_loc=_single_char(t[2], " "), res_name=t[3], chain_id=t[4], res_seq=res_seq,
                i_code=_single_char(t[6], " "), occupancy=parse_f64(t[10]),
                temp_factor=parse_f64(t[11]), element=element, is_hetero_atom=False)


@dataclass
class Structure:
    atoms: list
    models: list  # list of lists of (x, y, z)

    def count_models(self):
        return len(self.models)

    def count_chains(self):
        return len({a.chain_id for a in self.atoms})

    def count_residues(self):
        """structure.rs:236-239 (SameResidue over neighbouring atoms, + 1)"""
        n = 0
        for a, b in zip(self.atoms, self.atoms[1:]):
            if not (a.chain_id == b.chain_id and a.res_seq == b.res_seq and a.i_code == b.i_code):
                n += 1
        return n + 1


def structure_from_cif(text) -> Structure:
    """Deposit::from_cif_reader (load_cif.rs:33-40) + PdbAtom::from_cif_data."""
    blocks = read_cif(text)
    if len(blocks) > 1:
        raise IngestError("more than one data block")
    if not blocks:
        raise IngestError("no data block")
    atoms, models, ids = [], [], {}
    for t in atom_site_rows(blocks[0]):
        mid = parse_usize(t[13])
        if mid not in ids:
            ids[mid] = len(ids)
            models.append([])
        m = ids[mid]
        pos = (parse_f64(t[7]), parse_f64(t[8]), parse_f64(t[9]))
        models[m].append(pos)
        if len(ids) == 1:
            atoms.append(_cif_atom(t))
    return Structure(atoms, models)


def _slice(b: bytes, i: int, j: int | None = None) -> str:
    """&line[i..j] on a Rust str: byte offsets, which must lie on character
    boundaries and inside the line (else the reference panics)."""
    j = len(b) if j is None else j
    if j > len(b) or i > j:
        raise IngestError(f"PDB line too short: {b!r}")
    for k in (i, j):
        if k < len(b) and (b[k] & 0xC0) == 0x80:
            raise IngestError(f"PDB field splits a character: {b!r}")
    return b[i:j].decode("utf-8")


def _pdb_atom(line: str) -> Atom:
    """PdbAtom::from_atom_line, pdb_atom.rs:87-116."""
    b = line.encode()
    element = rust_trim(_slice(b, 77)) if len(b) >= 78 else None
    return Atom(serial=parse_i32(rust_trim(_slice(b, 6, 11))), name=_slice(b, 12, 16),
                alt_loc=_slice(b, 16, 17)[0], res_name=_slice(b, 17, 20), chain_id=_slice(b, 21, 22),
                res_seq=parse_i32(rust_trim(_slice(b, 22, 26))), i_code=_slice(b, 26, 27)[0],
                occupancy=parse_f64(rust_trim(_slice(b, 54, 60))),
                temp_factor=parse_f64(rust_trim(_slice(b, 60, 66))), element=element,
                is_hetero_atom=line.startswith("H"))


def _pdb_pos(line: str):
    """Vec3::from_pdb_line (bioshell-core/src/vec3.rs:241-246)"""
    b = line.encode()
    return tuple(parse_f64(rust_trim(_slice(b, a, a + 8))) for a in (30, 38, 46))


def structure_from_pdb(text) -> Structure:
    """Deposit::from_pdb_reader, load_pdb.rs:43-115 (coordinate records)."""
    atoms, models = [], [[]]
    model_id = 0
    for line in rust_lines(text):
        b = line.encode()
        if len(b) < 6:
            continue
        record = rust_trim(_slice(b, 0, 6))
        if record in ("ATOM", "HETATM"):
            if model_id == 0:
                a = _pdb_atom(line)
                models[0].append(_pdb_pos(line))
                atoms.append(a)
            else:
                if model_id >= len(models):
                    raise IngestError("coordinates after ENDMDL without a MODEL record")
                models[model_id].append(_pdb_pos(line))
        elif record == "ENDMDL":
            model_id += 1
        elif record == "MODEL":
            if len(models) == model_id:
                models.append([])
    return Structure(atoms, models)


def detect_format(text) -> str:
    """is_cif_file / is_pdb_file on the first non-empty line."""
    if isinstance(text, str):
        text = text.encode()
    for raw in text.split(b"\n"):
        try:
            line = (raw[:-1] if raw.endswith(b"\r") else raw).decode("utf-8")
        except UnicodeDecodeError:
            raise IngestError("stream did not contain valid UTF-8") from None
        if line != "":
            if line.startswith("data_"):
                return "cif"
            if any(line.startswith(s) for s in ("HEADER", "ATOM", "HETATM", "REMARK", "MODEL")):
                return "pdb"
            break
    raise IngestError("invalid file format")


def read_text(path: str) -> bytes:
    """open_file (bioshell-core/src/io/utils.rs:312-330): .gz is inflated."""
    with (gzip.open(path, "rb") if path.endswith(".gz") else open(path, "rb")) as f:
        return f.read()


def load_structure(text, fmt: str = "auto") -> Structure:
    if fmt == "auto":
        fmt = detect_format(text)
    return structure_from_cif(text) if fmt == "cif" else structure_from_pdb(text)


# ------------------------------------------------------------------ predicates

def format_name(name: str) -> str:
    """format_name! (pdb_atom_filters.rs:351-358): centred to 4 unless 4 long."""
    if len(name.encode()) == 4:
        return name
    if len(name) >= 4:
        return name
    left = (4 - len(name)) // 2
    return " " * left + name + " " * (4 - len(name) - left)


def is_hydrogen(a: Atom) -> bool:
    """pdb_atom.rs:352-365"""
    if a.element == "H":
        return True
    b = a.name.encode()
    if not b:
        return False
    if b[0:1] == b"H":
        return True
    if b[0:1].isdigit() or b[0:1] in (b" ", b"\t", b"\n", b"\x0c", b"\r"):
        return b[1:2] == b"H"
    return False


SELECTIONS = ("all", "ca", "cb", "not_hydrogen")


def selected(a: Atom, selection: str, chain: str | None) -> bool:
    if chain is not None and a.chain_id != chain:
        return False
    if selection == "all":
        return True
    if selection == "ca":
        return format_name(a.name) == " CA "
    if selection == "cb":
        return format_name(a.name) == " CB "
    if selection == "not_hydrogen":
        return not is_hydrogen(a)
    raise ValueError(selection)


@dataclass
class Ensemble:
    xyz: np.ndarray          # (n_models, n_selected, 3) float64
    atoms: list              # the selected atoms (metadata from the first model)
    n_structure_atoms: int
    n_chains: int
    n_residues: int


def ensemble(text, fmt: str = "auto", selection: str = "ca", chain: str | None = None) -> Ensemble:
    """Every model set in turn (structure.rs:277-285: a model whose atom count
    differs from the first model's is an error) and the atoms that pass the
    predicate kept, in file order."""
    s = load_structure(text, fmt)
    keep = [k for k, a in enumerate(s.atoms) if selected(a, selection, chain)]
    xyz = np.zeros((len(s.models), len(keep), 3))
    for m, coords in enumerate(s.models):
        if len(coords) != len(s.atoms):
            raise IngestError(f"wrong number of atoms in model {m}")
        for j, k in enumerate(keep):
            xyz[m, j] = coords[k]
    return Ensemble(xyz, [s.atoms[k] for k in keep], len(s.atoms), s.count_chains(),
                    s.count_residues())


# ------------------------------------------------------------------ PDB writing

def atom_line(a: Atom, x: float, y: float, z: float, chain_id: str | None = None) -> str:
    """impl Display for PdbAtom (pdb_atom.rs:256-278); charge is None."""
    rec = "HETATM" if a.is_hetero_atom else "ATOM  "
    elem = a.element if a.element is not None else "  "
    ch = a.chain_id if chain_id is None else chain_id
    return (f"{rec}{a.serial:>5} {a.name:^4}{a.alt_loc:>1}{a.res_name:>3} {ch:>1}"
            f"{a.res_seq:>4}{a.i_code:1}   {x:>8.3f}{y:>8.3f}{z:>8.3f}{a.occupancy:>6.2f}"
            f"{a.temp_factor:>6.2f}          {elem:>2}{'':2}")


def write_pdb(atoms, xyz_model) -> str:
    """write_pdb (structure.rs:828-850): chains in sorted order, renamed
    A..Z, 0..9 when any chain id is longer than one character."""
    chains = sorted({a.chain_id for a in atoms}, key=lambda c: c.encode())
    rename = any(len(c) > 1 for c in chains)
    codes = [chr(c) for c in range(ord("A"), ord("Z") + 1)] + [str(d) for d in range(10)]
    out = []
    for i, c in enumerate(chains):
        for k, a in enumerate(atoms):
            if a.chain_id == c:
                out.append(atom_line(a, *xyz_model[k], chain_id=codes[i] if rename else None) + "\n")
    return "".join(out)

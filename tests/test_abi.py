"""C-ABI surface and host-side logic of libgeauxdock_amd.so (CPU only: no
kernel is launched here)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import geauxdock_amd as g
from geauxdock_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in sorted(os.listdir(os.path.join(ROOT, "include")))
           if h.endswith(".h")]


def _declared(path):
    src = open(path).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    src = re.sub(r"//[^\n]*", "", src)
    names = set(re.findall(r"\b(gd_[a-z0-9_]+)\s*\(", src))
    return names


def test_library_loads():
    L = g.lib()
    assert L.gd_version().startswith(b"geauxdock_amd")


@pytest.mark.parametrize("header", HEADERS, ids=os.path.basename)
def test_every_declared_symbol_is_exported(header):
    L = C.CDLL(g.LIB_PATH)
    missing = [n for n in sorted(_declared(header)) if not hasattr(L, n)]
    assert not missing, missing


def test_dropin_symbols_exported():
    """The reference's own entry points (C++ linkage), for main_gpu/geauxdock.C."""
    out = os.popen("nm -D --defined-only %s" % g.LIB_PATH).read()
    for mangled in ["_Z9SetDevicev", "_Z4DockP7ComplexP6RecordPS0_PS2_PP5ParaTPP12gd_rng_state",
                    "_Z11DeviceAllocPP7ComplexPP6RecordPP12gd_rng_state",
                    "_Z10DeviceFreePP7ComplexPP6RecordPP12gd_rng_state",
                    "_Z8SetParaTP7ComplexPP5ParaT", "_Z7CopyH2DP7ComplexPS0_PP5ParaT",
                    "_Z10InitCurandPP12gd_rng_state"]:
        assert mangled in out, mangled


def test_python_prototypes_match_header():
    declared = set()
    for h in HEADERS:
        if h.endswith("geauxdock_amd.h"):
            declared |= _declared(h)
    assert declared == set(abi.PROTOTYPES), declared ^ set(abi.PROTOTYPES)


def test_missing_file_is_an_error_not_an_exit(ene, tmp_path):
    h = C.c_void_p()
    st = g.lib().gd_load_complex(b"/nonexistent.sdf", b"/nonexistent.ff", b"/nonexistent.pdb",
                                 b"x", C.byref(ene), 1, 0.044, 0.036, C.byref(h))
    assert st == abi.GD_ERR_IO


def test_unknown_ligand_type_rejected(ene, tmp_path):
    """An unknown SYBYL name (reference BADKDE = 50, an out-of-bounds table
    index there) rejects the job."""
    src = open(os.path.join(ROOT, "tests", "golden", "data", "1a07C1.sdf")).read()
    bad = src.replace("C.3 C.3 C.3 O.co2", "C.3 Xx C.3 O.co2", 1)
    p = tmp_path / "bad.sdf"
    p.write_text(bad)
    h = C.c_void_p()
    d = os.path.join(ROOT, "tests", "golden", "data")
    st = g.lib().gd_load_complex(str(p).encode(), os.path.join(d, "1a07C1-0.8.ff").encode(),
                                 os.path.join(d, "1a07C-1.pdb").encode(), b"1a07C1", C.byref(ene),
                                 1, 0.044, 0.036, C.byref(h))
    assert st == abi.GD_ERR_TYPE


def test_synthetic_sizes_and_determinism(ene):
    a = g.synth_complex(1200, 24, 5, 5, 1, seed=1, ene=ene)
    b = g.synth_complex(1200, 24, 5, 5, 1, seed=1, ene=ene)
    c = g.synth_complex(1200, 24, 5, 5, 1, seed=2, ene=ene)
    d = a.desc
    assert (d.prt_npoint, d.lig_natom, d.n_lig, d.n_prt, d.n_tmp) == (1200, 24, 5, 5, 1)
    assert d.kde_npoint == 120 * 24 and d.mcs_nrow == 64
    A, B, Cc = a.arrays(), b.arrays(), c.arrays()
    for k in ("prt_x", "lig_x", "kde_x", "mcs_x", "prt_t", "lig_t"):
        assert np.array_equal(A[k], B[k]), k
    assert not np.array_equal(A["prt_x"], Cc["prt_x"])
    # protein attributes follow the loader's typing rules
    assert abs(float(A["prt_ele"].sum())) < 1e-3
    cdc = (A["prt_c"] == 2) | ((A["prt_c"] == 0) & np.isin(A["prt_t"], [0]) & False)
    assert np.all(A["prt_cdc"][A["prt_c"] == 2] == 1)
    assert np.all(np.diff(A["prt_c"]) >= 0), "points sorted by class"
    assert np.all(np.diff(A["lig_t"]) >= 0), "atoms sorted by type"
    assert set(np.unique(A["lig_t"])) <= {2, 3, 4, 11, 13, 14, 16, 17, 18}
    # ligand conformers are centred on their own centroid
    assert np.abs(A["lig_x"].mean(1)).max() < 1e-4
    # every MCS row matches at least one atom
    assert np.all((A["mcs_x"] != np.float32(abi.GD_MCS_INVALID)).sum(1) >= 1)
    assert cdc is not None


def test_protein_typing_rules(golden_job):
    """Effective points of 1a07C-1.pdb: 104 residues -> 344 points by the
    loader's expansion (CA, PP from the third residue on, side chains)."""
    A = golden_job.arrays()
    t, c = A["prt_t"], A["prt_c"]
    assert (c == 0).sum() == 104            # one CA per residue
    assert (c == 1).sum() == 102            # PP for residues 3..104
    assert np.all(t[c == 0] == 0) and np.all(t[c == 1] == 1)
    assert np.all((t[c >= 2] >= 2) & (t[c >= 2] <= 29))
    assert np.all(A["prt_cdc"][c == 2] == 1) and np.all(A["prt_cdc"][c == 1] == 0)
    assert np.all(np.diff(c) >= 0)


def test_dropin_layout_matches_reference():
    """include/geauxdock_dropin.h reproduces sizeof/offsetof of the reference
    data model (tests/golden/ref_layout.json, written by make_ref_layout.py
    from the reference headers)."""
    import json
    import shutil
    import sys
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import make_ref_layout as m
    ref = json.load(open(os.path.join(ROOT, "tests", "golden", "ref_layout.json")))
    mine = m.measure('#include "%s"' % os.path.join(ROOT, "include", "geauxdock_dropin.h"), [])
    bad = {k: (mine.get(k), v) for k, v in ref.items() if mine.get(k) != v}
    assert not bad, bad
    assert len(ref) > 100


def test_host_side_argument_errors_and_csv_sink(tmp_path):
    """Entry points that touch no device: argument checks return GD_ERR_ARG,
    an unwritable record file GD_ERR_IO with a message, and the PrintRecord sink
    writes replica rep_lo..rep_hi (inclusive) of a dump's records, count[r]
    entries each, as "%3d %d" + " %.4f" x 10."""
    t = abi.gd_tiling()
    assert g.lib().gd_tiling_lookup(0, 3000, 1760, 0, C.byref(t)) == abi.GD_ERR_ARG
    assert g.lib().gd_tiling_lookup(40, 3000, 1760, 0, None) == abi.GD_ERR_ARG
    assert g.lib().gd_tiling_lookup(40, 3000, 1760, 1, C.byref(t)) == abi.GD_OK
    assert t.atoms_per_lane == 4 and 1 <= t.tile_points <= 3000 and t.waves_per_cu > 0
    assert g.lib().gd_job_set_streams(None, 2) == abi.GD_ERR_ARG
    w = C.c_void_p()
    assert g.lib().gd_record_csv_open(str(tmp_path / "x.txt").encode(), 3, 1, 1, C.byref(w)) == abi.GD_ERR_ARG
    assert g.lib().gd_record_csv_open(str(tmp_path / "no" / "x.txt").encode(), 0, 1, 1,
                                      C.byref(w)) == abi.GD_ERR_IO
    assert b"cannot open" in g.lib().gd_last_error()
    path = tmp_path / "rec.txt"
    assert g.lib().gd_record_csv_open(str(path).encode(), 1, 2, 1, C.byref(w)) == abi.GD_OK
    n_rep, cap = 4, 3
    rec = np.zeros((n_rep, cap, abi.REC_W), np.float32)
    for r in range(n_rep):
        for k in range(cap):
            rec[r, k, 0] = 10 * k + r
            rec[r, k, 7:17] = np.arange(10) * 0.5 + r + 0.01 * k
    count = np.array([3, 2, 0, 1], np.int32)
    g.lib().gd_record_csv_write(w, 0, 0, n_rep, cap, rec.ctypes.data_as(C.POINTER(C.c_float)),
                                count.ctypes.data_as(C.POINTER(C.c_int)))
    assert g.lib().gd_record_csv_close(w) == abi.GD_OK
    lines = path.read_text().splitlines()
    assert lines[0].split() == ["rep", "step", "vdw", "ele", "pmf", "psp", "hdb", "hpc", "kde", "lhm", "dst",
                                "total"]
    want = []
    for r in (1, 2):                      # rep_lo..rep_hi, with count[r] entries each
        for k in range(count[r]):
            want.append("%3d %d" % (r, int(rec[r, k, 0])) + "".join(" %.4f" % v for v in rec[r, k, 7:17]))
    assert lines[1:] == want


@pytest.mark.timeout(120)
def test_loader_and_generator_errors(ene, tmp_path):
    """Inputs the loaders and generators cannot use fail at once with a status
    and a message in gd_last_error(): a synthetic protein whose box has no
    room outside the pocket (the screen's fixed 14 A pocket, or a ligand wider
    than the box), a restraint record with missing columns (the message names
    its line), and a missing parameter file.  The smallest screen protein that
    fits still builds."""
    from tests.conftest import GOLDEN_DATA as d
    for make in (lambda: g.synth_screen(8, 1, 2, 6, 2, 2, 2, ene=ene),
                 lambda: g.synth_screen(80, 3, 20, 4, 2, 2, 2, ene=ene),
                 lambda: g.synth_complex(8, 80, 2, 2, 2, ene=ene)):
        with pytest.raises(g.GDError, match="no room outside a pocket"):
            make()
    job = g.synth_screen(100, 1, 12, 2, 2, 2, 2, ene=ene)
    assert job.desc.prt_npoint == 100
    ff = open(os.path.join(d, "1a07C1-0.8.ff")).read().splitlines()
    k = next(i for i, s in enumerate(ff) if s.startswith("KDE"))
    ff[k] = "KDE C.3 1.0"
    bad = tmp_path / "bad.ff"
    bad.write_text("\n".join(ff) + "\n")
    with pytest.raises(g.GDError, match="line %d: KDE record" % (k + 1)):
        g.load_complex(os.path.join(d, "1a07C1.sdf"), str(bad), os.path.join(d, "1a07C-1.pdb"), "1a07C1",
                       ene=ene)
    with pytest.raises(g.GDError, match="cannot open parameter file"):
        g.load_enepara(str(tmp_path))


_SETPARAT = r'''
import ctypes as C, json, struct, sys
sys.path.insert(0, sys.argv[1])
import geauxdock_amd as g
lay = json.load(open(sys.argv[2]))
f = getattr(g.lib(), "_Z8SetParaTP7ComplexPP5ParaT")
f.restype = None
cx = C.create_string_buffer(lay["Complex.sizeof"])
sizes = [int(v) for v in sys.argv[3].split(",")]   # n_lig n_prt n_tmp n_rep L P K M
C.memmove(C.addressof(cx) + lay["Complex.size"], struct.pack("8i", *sizes), 32)
pt = C.create_string_buffer(lay["ParaT.sizeof"])
f(cx, (C.c_void_p * 1)(C.addressof(pt)))
print("rep_end", struct.unpack_from("i", pt.raw, lay["ParaT.rep_end"])[0])
'''


@pytest.mark.parametrize("sizes,msg", [
    ("2,2,2,8,40,300,100,5", None),
    ("2,2,2,8,70,300,100,5", "ComplexSize.lig_natom = 70 is outside 0..64"),
    ("2,21,2,84,40,300,100,5", "ComplexSize.n_prt = 21 is outside 0..20"),
    ("2,2,2,9,40,300,100,5", "ComplexSize.n_rep = 9 is not n_lig x n_prt x n_tmp"),
    ("2,2,2,8,40,300,100,-1", "ComplexSize.mcs_nrow = -1 is outside 0..512"),
])
def test_dropin_sizes_beyond_capacity_fail_loudly(sizes, msg):
    """The drop-in reads the reference Complex's fixed-capacity arrays by its
    ComplexSize: sizes outside the capacities (or an n_rep that is not the
    ensemble product) end the program with a message instead of reading past
    the arrays (SetParaT touches no device)."""
    import subprocess
    import sys
    layout = os.path.join(ROOT, "tests", "golden", "ref_layout.json")
    r = subprocess.run([sys.executable, "-c", _SETPARAT, ROOT, layout, sizes], capture_output=True, text=True,
                       timeout=120)
    if msg is None:
        assert r.returncode == 0 and "rep_end 7" in r.stdout, r.stderr
    else:
        assert r.returncode == 1 and msg in r.stderr, (r.returncode, r.stderr)

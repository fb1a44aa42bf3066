"""The library's legacy flow-file writer and reader on the host (no GPU), against the oracle's
restatement of the format (oracle/flow_format.py, from src/eilmer/fluidblockio_old.d).

Pins: the variable lists of build_flow_variable_list for ideal air and 5-species air as
literal known answers; "%.18e" values that exercise signed zeros, subnormals, three-digit
exponents and the 19th significant digit; text and raw-binary files byte for byte equal to
the restatement's; the reader taking any variable order with extra variables, rejecting
a missing variable or a size mismatch as the reference throws; write-then-read returning
every double bit for bit.
"""
import gzip
import os

import numpy as np
import pytest

from gdtk_amd import hip
from oracle import flow_format as FF

AIR5 = ["N2", "O2", "NO", "N", "O"]

IDEAL_LIST = ["pos.x", "pos.y", "pos.z", "volume", "rho", "vel.x", "vel.y", "vel.z", "p", "a", "mu", "k",
              "mu_t", "k_t", "S", "massf[0]-air", "u", "T"]
AIR5_LIST = ["pos.x", "pos.y", "pos.z", "volume", "rho", "vel.x", "vel.y", "vel.z", "p", "a", "mu", "k",
             "mu_t", "k_t", "S", "massf[0]-N2", "massf[1]-O2", "massf[2]-NO", "massf[3]-N", "massf[4]-O",
             "dt_chem", "u", "T"]


def test_variable_lists_known_answers():
    assert FF.variable_list(["air"]) == IDEAL_LIST
    assert FF.variable_list(AIR5) == AIR5_LIST
    assert FF.variable_list(["my  gas"]) [15] == "massf[0]-my-gas"   # massfName squeezes white space
    assert hip.flow_nvars(1) == len(IDEAL_LIST)
    assert hip.flow_nvars(5) == len(AIR5_LIST)


def test_value_format_known_answers():
    assert FF.cell_line([1.0, -0.0, 0.0]) == \
        " 1.000000000000000000e+00 -0.000000000000000000e+00 0.000000000000000000e+00\n"
    # the exact binary values rounded to 19 significant digits (decimal.Decimal agrees)
    assert FF.cell_line([1.0e-300]) == " 1.000000000000000025e-300\n"
    assert FF.cell_line([5.0e-324]) == " 4.940656458412465442e-324\n"
    assert FF.cell_line([0.1]) == " 1.000000000000000056e-01\n"
    assert FF.cell_line([-123456.789]) == " -1.234567890000000043e+05\n"


def _records(ncell, nv, seed=3):
    rng = np.random.default_rng(seed)
    r = rng.standard_normal((ncell, nv)) * 10.0 ** rng.integers(-40, 40, (ncell, nv))
    r[0, :4] = [0.0, -0.0, 5.0e-324, 1.0e300]
    r[1, :3] = [np.nextafter(1.0, 2.0), np.nextafter(1.0, 0.0), 2.0 ** -1074 * 3]
    return r


@pytest.mark.parametrize("nsp,species", [(1, None), (5, AIR5)])
def test_text_file_matches_restatement(tmp_path, nsp, species):
    nic, njc, nkc, dims = 7, 5, 3, 3
    names = FF.variable_list(species or ["air"])
    rec = _records(nic * njc * nkc, len(names))
    fn = str(tmp_path / "b.flow.gz")
    hip.write_flow_records(fn, rec, dims, nic, njc, nkc, "gziptext", label="block-0", sim_time=1.2345e-4,
                           species=species)
    with gzip.open(fn, "rt", newline="") as f:
        got = f.read()
    assert got == FF.text(rec, names, "block-0", 1.2345e-4, dims, nic, njc, nkc)


def test_rawbinary_file_matches_restatement(tmp_path):
    nic, njc, nkc, dims = 6, 4, 1, 2
    names = FF.variable_list(AIR5)
    rec = _records(nic * njc, len(names), seed=5)
    fn = str(tmp_path / "b.flow")
    hip.write_flow_records(fn, rec, dims, nic, njc, nkc, "rawbinary", label="blk", sim_time=2.5, species=AIR5)
    assert open(fn, "rb").read() == FF.rawbinary(rec, names, "blk", 2.5, dims, nic, njc, nkc)


@pytest.mark.parametrize("fmt", ["gziptext", "rawbinary"])
def test_write_read_round_trip_bitwise(tmp_path, fmt):
    nic, njc, nkc = 9, 4, 2
    names = FF.variable_list(["air"])
    rec = _records(nic * njc * nkc, len(names), seed=7)
    fn = str(tmp_path / "b.flow")
    hip.write_flow_records(fn, rec, 3, nic, njc, nkc, fmt, sim_time=0.75)
    back, t = hip.read_flow_records(fn, 3, nic, njc, nkc, 1, fmt)
    assert t == 0.75
    assert np.array_equal(back.view(np.uint64), rec.view(np.uint64))


def test_reader_takes_any_variable_order(tmp_path):
    """A file whose variables come in another order, with extras, reads by name
    (scan_cell_data_from_variable_order_string)."""
    nic, njc, nkc = 5, 3, 1
    names = FF.variable_list(AIR5)
    rec = _records(nic * njc, len(names), seed=9)
    perm = np.random.default_rng(1).permutation(len(names))
    file_names = [names[i] for i in perm] + ["extra"]
    file_rec = np.concatenate([rec[:, perm], np.full((rec.shape[0], 1), 7.0)], axis=1)
    fn = str(tmp_path / "perm.flow.gz")
    FF.write_gziptext(fn, file_rec, file_names, "x", 3.0, 2, nic, njc, nkc)
    back, t = hip.read_flow_records(fn, 2, nic, njc, nkc, 5, "gziptext", species=AIR5)
    assert t == 3.0
    assert np.array_equal(back, rec)


def test_reader_errors(tmp_path):
    nic, njc, nkc = 4, 3, 1
    names = FF.variable_list(["air"])
    rec = _records(nic * njc, len(names), seed=2)
    fn = str(tmp_path / "e.flow.gz")
    FF.write_gziptext(fn, rec[:, 1:], names[1:], "x", 0.0, 2, nic, njc, nkc)   # no pos.x
    with pytest.raises(hip.HipError, match="Could not find variable: pos.x"):
        hip.read_flow_records(fn, 2, nic, njc, nkc, 1, "gziptext")
    FF.write_gziptext(fn, rec, names, "x", 0.0, 2, nic, njc, nkc)
    with pytest.raises(hip.HipError, match="mismatch in solution size"):
        hip.read_flow_records(fn, 2, nic + 1, njc, nkc, 1, "gziptext")
    with pytest.raises(hip.HipError, match="dimensions found"):
        hip.read_flow_records(fn, 3, nic, njc, nkc, 1, "gziptext")
    with pytest.raises(hip.HipError):
        hip.read_flow_records(os.path.join(str(tmp_path), "absent.gz"), 2, nic, njc, nkc, 1, "gziptext")
    # a cell line one value short
    txt = FF.text(rec, names, "x", 0.0, 2, nic, njc, nkc).split("\n")
    txt[9 + 5] = " ".join(txt[9 + 5].split(" ")[:-1])
    with gzip.open(fn, "wt", newline="") as f:
        f.write("\n".join(txt))
    with pytest.raises(hip.HipError, match="cell 5 has too few values"):
        hip.read_flow_records(fn, 2, nic, njc, nkc, 1, "gziptext")

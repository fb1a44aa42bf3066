"""Restatement of the reference's legacy flow-solution file format for a structured block.

TEST INFRASTRUCTURE ONLY: only tests/, __graft_entry__.smoke() and bench.py's CPU baseline
may import this module, as the checker of the library's flow files.

Follows src/eilmer/fluidblockio_old.d:
  - build_flow_variable_list (:107-151) and massfName (:98-102) for the configuration of
    the explicit path (no MHD, no quality, no radiation, laminar, one temperature, no
    local time stepping);
  - write_solution for an SFluidBlock (:388-447): the text header, one line per cell of
    " " + cell_data_as_string + "\\n" (:1487-1573, every value "%.18e"), gzip-compressed;
    or the raw binary layout (:395-418, cell_data_to_raw_binary :1577-1691);
  - read_legacy_solution for an SFluidBlock (:156-273): header checks, the variable list
    taken from the file (any order, extra variables ignored, a missing one an error).
Python's "%.18e" and float() are correctly rounded, as D's formattedWrite and parse are.
"""
import gzip
import re
import struct

import numpy as np


def variable_list(species):
    """The variable names of a cell record for a gas with the given species names."""
    v = ["pos.x", "pos.y", "pos.z", "volume", "rho", "vel.x", "vel.y", "vel.z",
         "p", "a", "mu", "k", "mu_t", "k_t", "S"]
    for i, name in enumerate(species):
        v.append(f"massf[{i}]-" + re.sub(r"[ \t]+", "-", name))
    if len(species) > 1:
        v.append("dt_chem")
    v += ["u", "T"]
    return v


def header_text(label, sim_time, names, dims, nic, njc, nkc):
    h = "structured_grid_flow 1.0\n"
    h += f"label: {label}\n"
    h += "sim_time: %.18e\n" % sim_time
    h += f"variables: {len(names)}\n"
    h += "".join(f' "{n}"' for n in names) + "\n"
    h += f"dimensions: {dims}\nnicell: {nic}\nnjcell: {njc}\nnkcell: {nkc}\n"
    return h


def cell_line(values):
    return " " + " ".join("%.18e" % float(x) for x in values) + "\n"


def text(records, names, label="", sim_time=0.0, dims=2, nic=1, njc=1, nkc=1):
    """The uncompressed text of a gziptext flow file."""
    return header_text(label, sim_time, names, dims, nic, njc, nkc) + "".join(cell_line(r) for r in records)


def write_gziptext(path, records, names, label="", sim_time=0.0, dims=2, nic=1, njc=1, nkc=1):
    with gzip.open(path, "wt", newline="") as f:
        f.write(text(records, names, label, sim_time, dims, nic, njc, nkc))


def rawbinary(records, names, label="", sim_time=0.0, dims=2, nic=1, njc=1, nkc=1):
    """The bytes of a rawbinary flow file (native little-endian ints and doubles)."""
    out = bytearray(b"structured_grid_flow 1.0")
    lab = label.encode()
    out += struct.pack("<i", len(lab)) + lab
    out += struct.pack("<d", sim_time)
    out += struct.pack("<i", len(names))
    for n in names:
        nb = n.encode()
        out += struct.pack("<i", len(nb)) + nb
    out += struct.pack("<4i", dims, nic, njc, nkc)
    out += np.ascontiguousarray(records, dtype="<f8").tobytes()
    return bytes(out)


def read_gziptext(path, want_names, dims, nic, njc, nkc):
    """(records in the order of want_names, sim_time, label) of a gziptext file."""
    with gzip.open(path, "rt", newline="") as f:
        lines = f.read().split("\n")
    version = lines[0].split("structured_grid_flow ", 1)[1]
    if version != "1.0":
        raise ValueError("format version " + version)
    label = lines[1].split("label:", 1)[1].strip()
    sim_time = float(lines[2].split("sim_time:", 1)[1])
    nvar = int(lines[3].split("variables:", 1)[1])
    names = [s.strip('"') for s in lines[4].split()]
    assert len(names) == nvar
    cols = []
    for n in want_names:
        if n not in names:
            raise ValueError("Could not find variable: " + n)
        cols.append(names.index(n))
    d = int(lines[5].split("dimensions:", 1)[1])
    ni = int(lines[6].split("nicell:", 1)[1])
    nj = int(lines[7].split("njcell:", 1)[1])
    nk = int(lines[8].split("nkcell:", 1)[1])
    if d != dims or ni != nic or nj != njc or nk != (nkc if dims == 3 else 1):
        raise ValueError("mismatch in solution size")
    ncell = nic * njc * nkc
    rows = np.array([[float(x) for x in lines[9 + c].split()] for c in range(ncell)])
    return rows[:, cols], sim_time, label


def read_rawbinary(path, want_names, ncell):
    """(records in the order of want_names, sim_time) of a rawbinary file
    (read_legacy_solution's raw binary branch for an SFluidBlock, fluidblockio_old.d:172-213)."""
    data = open(path, "rb").read()
    p = len(b"structured_grid_flow 1.0")
    if data[:p] != b"structured_grid_flow 1.0":
        raise ValueError("not a structured_grid_flow 1.0 file")
    lab = struct.unpack_from("<i", data, p)[0]
    p += 4 + lab
    sim_time = struct.unpack_from("<d", data, p)[0]
    p += 8
    nvar = struct.unpack_from("<i", data, p)[0]
    p += 4
    names = []
    for _ in range(nvar):
        ln = struct.unpack_from("<i", data, p)[0]
        names.append(data[p + 4:p + 4 + ln].decode())
        p += 4 + ln
    p += 16  # dims, nicell, njcell, nkcell
    rows = np.frombuffer(data[p:], "<f8").reshape(ncell, nvar)
    cols = []
    for n in want_names:
        if n not in names:
            raise ValueError("Could not find variable: " + n)
        cols.append(names.index(n))
    return rows[:, cols], sim_time

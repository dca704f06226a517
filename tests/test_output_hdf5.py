"""HDF5 block output and restart input on host blocks (no GPU).

What pins the files to the reference's:
* block names: the example of the reference's own design notes
  (doc/source/design/design-input.rst:276, "B00:0_00:0_00:0"), names worked
  by hand from Block::name / Index::bit_string (src/Cello/mesh_Block.cpp:
  798-821, src/Cello/mesh_Index.cpp:544-575), uniqueness over a whole tree and
  decoding a name back to its block;
* the "index" attribute: the three words of Cello's NodeBits union laid out
  by ctypes from the same bit-field declaration (src/Cello/mesh_Index.hpp:
  16-40) and filled the way Index::set_array / set_level / set_child do
  (src/Cello/mesh_Index.cpp:340-411);
* file names: C printf semantics of Output::expand_name_
  (src/Cello/io_Output.cpp:125-207) against Python's %-formatting;
* the layout: the file is read back through libhdf5's C API directly
  (tests/h5read.py shares no code with the writer): the attribute names,
  types and shapes of OutputData / IoSimulation / IoHierarchy / IoBlock /
  IoEnzoBlock, their values in closed form, datasets bit for bit, the
  .block_list / .file_list text files; and with the HDF5 tools' h5diff the
  way the reference's checkpoint-restart test compares block groups
  (tools/ckpt_restart_test.py:200-211).
"""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

import h5read

pytestmark = pytest.mark.skipif(not h5read.available(), reason="libhdf5 not present")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
H5DIFF = "/opt/conda/bin/h5diff"
ERR_CONFIG, ERR_ARG, ERR_IO = 1, 2, 6


def _io():
    import enzoe_io
    return enzoe_io


# ------------------------------------------------------------------ names

@pytest.mark.parametrize("level,gi,rb,rank,name", [
    (1, (0, 0, 0), (4, 4, 4), 3, "B00:0_00:0_00:0"),  # design-input.rst:276
    (0, (0, 0, 0), (1, 1, 1), 3, "B__"),
    (0, (1, 0, 0), (2, 2, 1), 3, "B1_0_"),
    (0, (1, 2, 3), (4, 4, 4), 3, "B01_10_11"),
    (0, (2, 0, 0), (3, 1, 1), 3, "B10__"),
    (1, (3, 0, 1), (2, 2, 2), 3, "B1:1_0:0_0:1"),
    (2, (2, 1, 3), (1, 1, 1), 3, "B:10_:01_:11"),
    (0, (1, 1, 0), (2, 2, 1), 2, "B1_1"),
    (0, (5, 0, 0), (8, 1, 1), 1, "B101"),
])
def test_block_names(level, gi, rb, rank, name):
    assert _io().block_name(level, gi, rb, rank) == name


def _decode(name, rb):
    """Block position and level from a name (the inverse of Block::name)."""
    axes = name[1:].split("_")
    level, gi = 0, []
    for a, s in enumerate(axes):
        root, _, child = s.partition(":")
        level = len(child)
        r = int(root, 2) if root else 0
        c = int(child, 2) if child else 0
        gi.append((r << level) | c)
    return level, tuple(gi)


def test_block_names_are_unique_and_decode_back():
    io = _io()
    rb = (3, 2, 1)
    seen = set()
    for level in range(3):
        for z in range(rb[2] << level):
            for y in range(rb[1] << level):
                for x in range(rb[0] << level):
                    n = io.block_name(level, (x, y, z), rb)
                    assert n not in seen
                    seen.add(n)
                    assert _decode(n, rb) == (level, (x, y, z))
    assert len(seen) == 6 * (1 + 8 + 64)


class NodeBits(ctypes.Structure):
    """src/Cello/mesh_Index.hpp:16-40, laid out by ctypes."""
    _fields_ = [("level", ctypes.c_uint, 2), ("array", ctypes.c_uint, 10),
                ("tree", ctypes.c_uint, 20)]


class IndexUnion(ctypes.Union):
    _fields_ = [("a", NodeBits * 3), ("v", ctypes.c_int * 3)]


def _reference_index(level, gi):
    """An Index built as Block::index_from_global builds it
    (mesh_Block.cpp:936-950): set_array, set_level, then set_child per level
    from the finest up, each through the bit fields (which truncate)."""
    u = IndexUnion()
    for a in range(3):
        u.a[a].array = gi[a] >> level
    shift = 2  # INDEX_BITS_LEVEL
    mask = ~(1 << shift)
    u.a[0].level = ((level >> 0) & mask) & 3
    u.a[1].level = ((level >> shift) & mask) & 3
    u.a[2].level = ((level >> (2 * shift - 1)) & (mask - 1)) & 3
    for i in range(level):
        lv = level - i
        s = 20 - lv
        for a in range(3):
            bit = (gi[a] >> i) & 1
            u.a[a].tree = ((u.a[a].tree & ~(1 << s)) | (bit << s)) & ((1 << 20) - 1)
    return u


@pytest.mark.parametrize("level,gi,rb", [
    (0, (0, 0, 0), (1, 1, 1)), (0, (3, 1, 2), (4, 2, 3)), (1, (7, 2, 5), (4, 2, 3)),
    (3, (17, 9, 30), (4, 2, 4)), (5, (100, 33, 63), (4, 2, 2)),
    (20, ((1 << 20) + 12345, 777, 3), (2, 1, 1))])
def test_index_words_match_the_bitfield_layout(level, gi, rb):
    u = _reference_index(level, gi)
    assert _io().block_index_words(level, gi, rb) == tuple(u.v)
    # Index::level (mesh_Index.cpp:327-336) reads the level back
    lx, ly, lz = u.a[0].level, u.a[1].level, u.a[2].level >> 1
    assert lx + 4 * (ly + 4 * lz) == level
    # Block::index_global (mesh_Block.cpp:902-932) reads the position back
    for a in range(3):
        g = u.a[a].array
        for lv in range(1, level + 1):
            g = (g << 1) | ((u.a[a].tree >> (20 - lv)) & 1)
        assert g == gi[a]


def test_block_name_rejects_out_of_range():
    io = _io()
    for level, gi, rb in [(0, (2, 0, 0), (2, 1, 1)), (-1, (0, 0, 0), (1, 1, 1)),
                          (21, (0, 0, 0), (1, 1, 1)), (0, (0, 0, 0), (1025, 1, 1))]:
        with pytest.raises(io.EnzoeIoError) as e:
            io.block_name(level, gi, rb)
        assert e.value.status == ERR_ARG


# ------------------------------------------------------------------ file names

@pytest.mark.parametrize("fmt,args,kw,py", [
    ("data-%02d.h5", ["cycle"], dict(cycle=7), "data-%02d.h5" % 7),
    ("Dir_%04d", ["cycle"], dict(cycle=123), "Dir_%04d" % 123),
    ("out-%d-%d.h5", ["cycle", "proc"], dict(cycle=3, proc=1), "out-%d-%d.h5" % (3, 1)),
    ("t-%5.2f-%03d", ["time", "count"], dict(time=0.25, count=4), "t-%5.2f-%03d" % (0.25, 4)),
    ("t%.3e.h5", ["time"], dict(time=12.5), "t%.3e.h5" % 12.5),
    ("f-%d", ["flipflop"], dict(count=3), "f-1"),
    ("plain.h5", [], {}, "plain.h5"),
])
def test_expand_name(fmt, args, kw, py):
    assert _io().expand_name(fmt, args, **kw) == py


@pytest.mark.parametrize("fmt,args", [
    ("data-%d-%d.h5", ["cycle"]),   # conversions != variables
    ("data.h5", ["cycle"]),
    ("data-%d.h5", ["step"]),       # unknown variable
    ("data-%f.h5", ["cycle"]),      # conversion of the wrong type
    ("data-%d.h5", ["time"]),
    ("data-%s.h5", ["cycle"]),
    ("data-%ld.h5", ["cycle"]),
    ("data-\\%d.h5", ["cycle"]),     # escaped percent (io_Output.cpp:143-146)
])
def test_expand_name_rejects(fmt, args):
    io = _io()
    with pytest.raises(io.EnzoeIoError) as e:
        io.expand_name(fmt, args, cycle=1)
    assert e.value.status == ERR_CONFIG


# ------------------------------------------------------------------ files

def _host_grid(nb, n, g, seed=5, mhd=True):
    """Blocks of a uniform grid with random fields (x fastest)."""
    io = _io()
    rng = np.random.default_rng(seed)
    m = [n[a] + 2 * g for a in range(3)]
    blocks = []
    for _ in range(nb[0] * nb[1] * nb[2]):
        arrs = {f: rng.standard_normal((m[2], m[1], m[0]))
                for f in ("density", "velocity_x", "velocity_y", "velocity_z",
                          "total_energy")}
        if mhd:
            for c in "xyz":
                arrs["bfield_" + c] = rng.standard_normal((m[2], m[1], m[0]))
            arrs["bfieldi_x"] = rng.standard_normal((m[2], m[1], m[0] + 1))
            arrs["bfieldi_y"] = rng.standard_normal((m[2], m[1] + 1, m[0]))
            arrs["bfieldi_z"] = rng.standard_normal((m[2] + 1, m[1], m[0]))
        blocks.append(io.HostBlock(arrs, g))
    return blocks


BLOCK_ATTRS = ["num_field_data", "index", "lower", "upper", "cycle", "time", "dt", "array",
               "index_order", "count_order", "enzo_dt", "enzo_GridLeftEdge",
               "enzo_GridDimension", "enzo_GridStartIndex", "enzo_GridEndIndex",
               "enzo_CellWidth", "enzo_redshift"]
INT, FLOAT = h5read.H5T_INTEGER, h5read.H5T_FLOAT
BLOCK_TYPES = {"num_field_data": (INT, 4, (1,)), "index": (INT, 4, (3,)),
               "lower": (FLOAT, 8, (3,)), "upper": (FLOAT, 8, (3,)), "cycle": (INT, 4, (1,)),
               "time": (FLOAT, 8, (1,)), "dt": (FLOAT, 8, (1,)), "array": (INT, 4, (3,)),
               "index_order": (INT, 8, (1,)), "count_order": (INT, 8, (1,)),
               "enzo_dt": (FLOAT, 8, (1,)), "enzo_GridLeftEdge": (FLOAT, 8, (3,)),
               "enzo_GridDimension": (INT, 4, (3,)), "enzo_GridStartIndex": (INT, 4, (3,)),
               "enzo_GridEndIndex": (INT, 4, (3,)), "enzo_CellWidth": (FLOAT, 8, (3,)),
               "enzo_redshift": (FLOAT, 8, (1,))}


@pytest.fixture
def in_tmp(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    return tmp_path


def test_output_layout(in_tmp):
    io = _io()
    nb, n, g = (2, 2, 1), (8, 6, 4), 3
    blocks = _host_grid(nb, n, g)
    fields = ["density", "velocity_x", "total_energy", "bfield_z", "bfieldi_x",
              "bfieldi_y", "bfieldi_z"]
    out = io.OutputData(fields, name=("data-%03d.h5", "cycle"), dir=("Dir_%03d", "cycle"))
    path = out.write(blocks, (0.0, 0.0, 0.0), (2.0, 1.0, 0.5), nb, cycle=5, time=0.125,
                     dt=0.01)
    assert path == "Dir_005/data-005.h5"
    # lists (io_OutputData.cpp:154-193)
    names = ["B0_0_", "B1_0_", "B0_1_", "B1_1_"]
    with open("Dir_005/Dir_005.block_list") as f:
        assert f.read() == "".join("%s data-005.h5\n" % b for b in names)
    with open("Dir_005/Dir_005.file_list") as f:
        assert f.read() == "data-005.h5\n"
    assert io.read_block_list("Dir_005/Dir_005.block_list") == {b: "data-005.h5" for b in names}

    with h5read.H5File(path) as h:
        assert sorted(h.groups("/")) == sorted(names)
        root = h.attrs("/")
        assert sorted(root) == sorted(["rank", "cycle", "time", "dt", "lower", "upper",
                                       "max_level", "version"])
        assert root["rank"].tolist() == [3] and root["cycle"].tolist() == [5]
        assert root["time"].tolist() == [0.125] and root["dt"].tolist() == [0.01]
        assert root["lower"].tolist() == [0, 0, 0] and root["upper"].tolist() == [2, 1, 0.5]
        assert root["max_level"].tolist() == [0]
        assert bytes(root["version"].astype(np.uint8)) == b"1.0.0"
        types = h.attr_types("/")
        assert types["version"] == (INT, 1, (5,)) and types["rank"] == (INT, 4, (1,))
        for i, b in enumerate(names):
            bx, by = i % 2, i // 2
            gp = "/" + b
            assert sorted(h.attr_types(gp).items()) == sorted(BLOCK_TYPES.items())
            a = h.attrs(gp)
            assert a["num_field_data"].tolist() == [1]
            assert a["index"].tolist() == [bx << 2, by << 2, 0]
            assert a["array"].tolist() == [bx, by, 0]
            assert a["lower"].tolist() == [1.0 * bx, 0.5 * by, 0.0]
            assert a["upper"].tolist() == [1.0 * (bx + 1), 0.5 * (by + 1), 0.5]
            assert a["cycle"].tolist() == [5] and a["time"].tolist() == [0.125]
            assert a["dt"].tolist() == [0.01] and a["enzo_dt"].tolist() == [0.01]
            assert a["index_order"].tolist() == [0] and a["count_order"].tolist() == [0]
            assert a["enzo_GridLeftEdge"].tolist() == a["lower"].tolist()
            assert a["enzo_GridDimension"].tolist() == [14, 12, 10]
            assert a["enzo_GridStartIndex"].tolist() == [3, 3, 3]
            assert a["enzo_GridEndIndex"].tolist() == [10, 8, 6]
            assert a["enzo_CellWidth"].tolist() == [1.0 / 8, 0.5 / 6, 0.5 / 4]
            assert a["enzo_redshift"].tolist() == [0.0]
            for f in fields:
                d = h.dataset(gp + "/field_" + f)
                assert d.dtype == np.float64 and d.shape == blocks[i].shapes[f]
                assert np.array_equal(d, blocks[i].arrays[f]), (b, f)


def test_output_in_working_directory_and_round_trip(in_tmp):
    io = _io()
    nb, n, g = (1, 1, 1), (6, 5, 4), 3
    blocks = _host_grid(nb, n, g, seed=9, mhd=False)
    out = io.OutputData(None, name=("hydro-%d-%05.3f.h5", "cycle", "time"))
    path = out.write(blocks, (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0), nb, cycle=2, time=0.5,
                     dt=0.25)
    assert path == "./hydro-2-0.500.h5"
    # no directory: lists named after the file without its extension
    assert open("hydro-2-0.500.block_list").read() == "B__ hydro-2-0.500.h5\n"
    assert io.read_state(path) == (2, 0.5, 0.25)
    back = io.HostBlock({k: np.zeros_like(v) for k, v in blocks[0].arrays.items()}, g)
    assert io.read_block(path, "B__", back) == (2, 0.5, 0.25)
    for k, v in blocks[0].arrays.items():
        assert np.array_equal(back.arrays[k], v), k


@pytest.mark.skipif(not os.path.exists(H5DIFF), reason="h5diff not present")
def test_h5diff_compares_block_groups_like_the_checkpoint_restart_test(in_tmp):
    io = _io()
    nb, n, g = (2, 1, 1), (4, 4, 4), 3
    a = _host_grid(nb, n, g, seed=1)
    out = io.OutputData(None, name=("a.h5",))
    pa = out.write(a, (0, 0, 0), (1, 1, 1), nb, 0, 0.0, 0.1)
    pb = io.OutputData(None, name=("b.h5",)).write(a, (0, 0, 0), (1, 1, 1), nb, 0, 0.0, 0.1)
    for blk in ("B0__", "B1__"):
        r = subprocess.run([H5DIFF, "-c", pa, pb, "/" + blk], capture_output=True)
        assert r.returncode == 0 and r.stdout == b"", r.stdout
    a[1].arrays["bfieldi_y"][2, 3, 1] += 1e-12
    pc = io.OutputData(None, name=("c.h5",)).write(a, (0, 0, 0), (1, 1, 1), nb, 0, 0.0, 0.1)
    assert subprocess.run([H5DIFF, "-c", pa, pc, "/B0__"], capture_output=True).returncode == 0
    assert subprocess.run([H5DIFF, "-c", pa, pc, "/B1__"], capture_output=True).returncode == 1


def test_output_and_input_errors(in_tmp):
    io = _io()
    nb, n, g = (1, 1, 1), (4, 4, 4), 3
    blocks = _host_grid(nb, n, g, mhd=False)
    out = io.OutputData(None, name=("e.h5",))
    path = out.write(blocks, (0, 0, 0), (1, 1, 1), nb, 0, 0.0, 0.1)
    with pytest.raises(io.EnzoeIoError) as e:            # no such block
        io.read_block(path, "B1__", blocks[0])
    assert e.value.status == ERR_IO
    small = io.HostBlock({"density": np.zeros((9, 10, 10))}, g)
    with pytest.raises(io.EnzoeIoError) as e:            # extents differ
        io.read_block(path, "B__", small)
    assert e.value.status == ERR_IO
    other = io.HostBlock({"pressure": np.zeros((10, 10, 10))}, g)
    with pytest.raises(io.EnzoeIoError) as e:            # no such dataset
        io.read_block(path, "B__", other)
    assert e.value.status == ERR_IO
    with pytest.raises(io.EnzoeIoError) as e:            # no such file
        io.read_state("missing.h5")
    assert e.value.status == ERR_IO
    with pytest.raises(io.EnzoeIoError) as e:            # block outside the grid
        out.write(blocks, (0, 0, 0), (1, 1, 1), nb, 0, 0.0, 0.1, indices=[(1, 0, 0)])
    assert e.value.status == ERR_ARG
    with pytest.raises(io.EnzoeIoError) as e:            # the same block twice
        out.write(blocks * 2, (0, 0, 0), (1, 1, 1), nb, 0, 0.0, 0.1,
                  indices=[(0, 0, 0), (0, 0, 0)])
    assert e.value.status == ERR_IO
    with pytest.raises(io.EnzoeIoError) as e:            # bad name variables
        io.OutputData(None, name=("x-%d.h5", "step")).write(blocks, (0, 0, 0), (1, 1, 1),
                                                            nb, 0, 0.0, 0.1)
    assert e.value.status == ERR_CONFIG


def test_fields_larger_than_a_slab_and_two_dimensional_blocks(in_tmp):
    """A 37 MB field goes to the file in two slabs of planes; a 2D block's
    fields are (my, mx) datasets (io_OutputData.cpp:240-246) and its z
    extents are those of an axis without ghost zones."""
    io = _io()
    rng = np.random.default_rng(3)
    big = io.HostBlock({"density": rng.standard_normal((70, 256, 257))}, 3)
    p = io.OutputData(None, name=("big.h5",)).write([big], (0, 0, 0), (1, 1, 1), (1, 1, 1),
                                                    0, 0.0, 0.0)
    with h5read.H5File(p) as h:
        assert np.array_equal(h.dataset("/B__/field_density"), big.arrays["density"])
    back = io.HostBlock({"density": np.zeros((70, 256, 257))}, 3)
    io.read_block(p, "B__", back)
    assert np.array_equal(back.arrays["density"], big.arrays["density"])

    flat = [io.HostBlock({"density": rng.standard_normal((1, 10, 12)),
                          "bfieldi_y": rng.standard_normal((1, 11, 12))}, 3)
            for _ in range(2)]
    p = io.OutputData(None, name=("flat.h5",)).write(flat, (0, 0, 0), (2, 1, 1), (2, 1, 1),
                                                     0, 0.0, 0.0, rank=2)
    assert open("flat.block_list").read() == "B0_ flat.h5\nB1_ flat.h5\n"
    with h5read.H5File(p) as h:
        a = h.attrs("/B1_")
        assert a["enzo_GridDimension"].tolist() == [12, 10, 1]
        assert a["enzo_GridStartIndex"].tolist() == [3, 3, 0]
        assert a["enzo_GridEndIndex"].tolist() == [8, 6, 0]
        assert a["enzo_CellWidth"].tolist() == [1.0 / 6, 1.0 / 4, 1.0]
        assert a["lower"].tolist() == [1.0, 0.0, 0.0]
        d = h.dataset("/B1_/field_bfieldi_y")
        assert d.shape == (11, 12)
        assert np.array_equal(d, flat[1].arrays["bfieldi_y"][0])

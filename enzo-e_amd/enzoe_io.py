"""Python binding of the HDF5 block output and restart input
(``libenzoe_io.so``, C ABI in ``include/enzoe_output.h``).

Files follow the layout Cello's OutputData writes (src/Cello/io_OutputData.cpp:
68-250) and IoEnzoReader reads back (src/Enzo/io/control_restart.cpp:559-720):
one HDF5 file per output with the simulation attributes at the root, one group
``/<block name>`` per block holding its IoBlock/IoEnzoBlock attributes and one
``field_<name>`` dataset per field (ghost zones included), plus the
``.block_list`` and ``.file_list`` text files.

Blocks are any object with ``shapes`` ({field: (mz, my, mx)}), ``ptr(name)``,
``ghost`` and ``on_device``: a ``enzoe_vlct.DeviceBlock`` (device memory,
copied through pinned buffers) or a ``HostBlock`` (numpy arrays).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
IO_LIB_PATH = os.path.join(HERE, "libenzoe_io.so")

ENZOE_ERR_IO = 6


class EnzoeIoError(RuntimeError):
    def __init__(self, what, rc):
        super().__init__("%s returned status %d" % (what, rc))
        self.status = rc


class BlockIndex(ctypes.Structure):
    _fields_ = [("level", ctypes.c_int), ("global_", ctypes.c_int * 3)]


class OutputField(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("data", ctypes.c_void_p),
                ("mx", ctypes.c_int), ("my", ctypes.c_int), ("mz", ctypes.c_int)]


class OutputHierarchy(ctypes.Structure):
    _fields_ = [("rank", ctypes.c_int), ("lower", ctypes.c_double * 3),
                ("upper", ctypes.c_double * 3), ("root_blocks", ctypes.c_int * 3),
                ("max_level", ctypes.c_int), ("cycle", ctypes.c_int),
                ("time", ctypes.c_double), ("dt", ctypes.c_double)]


class OutputBlock(ctypes.Structure):
    _fields_ = [("index", BlockIndex), ("cycle", ctypes.c_int), ("time", ctypes.c_double),
                ("dt", ctypes.c_double), ("n", ctypes.c_int * 3), ("ghost", ctypes.c_int * 3),
                ("fields", ctypes.POINTER(OutputField)), ("nfields", ctypes.c_int),
                ("on_device", ctypes.c_int)]


_io = None


def io_lib():
    """libenzoe_io.so; raises if it was not built."""
    global _io
    if _io is None:
        if not os.path.exists(IO_LIB_PATH):
            raise RuntimeError("libenzoe_io.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(IO_LIB_PATH)
        vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        PI, PD = ctypes.POINTER(ci), ctypes.POINTER(cd)
        strs = ctypes.POINTER(ctypes.c_char_p)
        L.enzoe_block_name.argtypes = [ctypes.POINTER(BlockIndex), PI, ci, ctypes.c_char_p, ci]
        L.enzoe_block_index_words.argtypes = [ctypes.POINTER(BlockIndex), PI, PI]
        L.enzoe_output_expand_name.argtypes = [ctypes.c_char_p, strs, ci, ci, cd, ci, ci,
                                               ctypes.c_char_p, ci]
        L.enzoe_output_open.argtypes = [ctypes.c_char_p, strs, ci, ctypes.c_char_p, strs, ci,
                                        ci, ci, ctypes.POINTER(OutputHierarchy),
                                        ctypes.POINTER(vp)]
        L.enzoe_output_write_block.argtypes = [vp, ctypes.POINTER(OutputBlock), vp]
        L.enzoe_output_file_path.argtypes = [vp]
        L.enzoe_output_file_path.restype = ctypes.c_char_p
        L.enzoe_output_close.argtypes = [vp]
        L.enzoe_input_read_state.argtypes = [ctypes.c_char_p, PI, PD, PD]
        L.enzoe_input_read_block.argtypes = [ctypes.c_char_p, ctypes.c_char_p,
                                             ctypes.POINTER(OutputField), ci, ci, vp, PI, PD,
                                             PD]
        _io = L
    return _io


def _check(rc, what):
    if rc != 0:
        raise EnzoeIoError(what, rc)


def _index(level, global_index):
    return BlockIndex(int(level), (ctypes.c_int * 3)(*[int(g) for g in global_index]))


def _ints3(v):
    return (ctypes.c_int * 3)(*[int(x) for x in v])


def _strs(items):
    items = list(items)
    return (ctypes.c_char_p * max(1, len(items)))(*[s.encode() for s in items]), len(items)


def block_name(level, global_index, root_blocks, rank=3):
    """Block::name of the block at `global_index` on refinement `level`."""
    buf = ctypes.create_string_buffer(256)
    _check(io_lib().enzoe_block_name(ctypes.byref(_index(level, global_index)),
                                     _ints3(root_blocks), int(rank), buf, 256),
           "enzoe_block_name")
    return buf.value.decode()


def block_index_words(level, global_index, root_blocks):
    """The three packed words of the block's Cello Index."""
    w = (ctypes.c_int * 3)()
    _check(io_lib().enzoe_block_index_words(ctypes.byref(_index(level, global_index)),
                                            _ints3(root_blocks), w),
           "enzoe_block_index_words")
    return tuple(w)


def expand_name(fmt, args=(), cycle=0, time=0.0, count=0, proc=0):
    """Output::expand_name_ of a file or directory name."""
    a, n = _strs(args)
    buf = ctypes.create_string_buffer(1024)
    _check(io_lib().enzoe_output_expand_name(fmt.encode(), a, n, int(cycle), float(time),
                                             int(count), int(proc), buf, 1024),
           "enzoe_output_expand_name")
    return buf.value.decode()


class HostBlock:
    """A block's fields in host memory (numpy arrays, C order (z, y, x))."""

    on_device = False

    def __init__(self, arrays, ghost):
        self.arrays = {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in arrays.items()}
        self.shapes = {k: v.shape for k, v in self.arrays.items()}
        self.ghost = ghost

    def ptr(self, name):
        return self.arrays[name].ctypes.data


def _fields(block, field_list):
    names = list(field_list) if field_list is not None else list(block.shapes)
    keep = [n.encode() for n in names]  # the C strings must outlive the call
    arr = (OutputField * max(1, len(names)))()
    for i, n in enumerate(names):
        if n not in block.shapes:
            raise KeyError("block has no field %r" % n)
        mz, my, mx = block.shapes[n]
        arr[i] = OutputField(keep[i], block.ptr(n), mx, my, mz)
    return arr, len(names), keep


def _active(block, rank):
    """Active cells per axis from the density field's extents."""
    mz, my, mx = block.shapes["density"]
    g = block.ghost
    m = (mx, my, mz)
    return [m[a] - 2 * g if a < rank else 1 for a in range(3)], \
        [g if a < rank else 0 for a in range(3)]


class OutputData:
    """A Cello "data" output file group: Output:<group>:name / dir (a format
    and the variables it takes, io_Output.cpp:125-207) and field_list."""

    def __init__(self, field_list=None, name=("data-%04d.h5", "cycle"), dir=None):
        self.field_list = field_list
        self.name = (name[0], list(name[1:]))
        self.dir = (dir[0], list(dir[1:])) if dir else ("", [])

    def write(self, blocks, lower, upper, root_blocks, cycle, time, dt, indices=None,
              levels=None, count=0, proc=0, rank=3, max_level=0, stream=None):
        """Writes one file holding `blocks` (a uniform root-level grid, x
        fastest, unless `indices`/`levels` are given) and its block and file
        lists.  Returns the path of the HDF5 file."""
        L = io_lib()
        h = OutputHierarchy(int(rank), (ctypes.c_double * 3)(*lower),
                            (ctypes.c_double * 3)(*upper), _ints3(root_blocks),
                            int(max_level), int(cycle), float(time), float(dt))
        fa, nfa = _strs(self.name[1])
        da, nda = _strs(self.dir[1])
        handle = ctypes.c_void_p()
        _check(L.enzoe_output_open(self.name[0].encode(), fa, nfa,
                                   self.dir[0].encode() if self.dir[0] else None, da, nda,
                                   int(count), int(proc), ctypes.byref(h),
                                   ctypes.byref(handle)), "enzoe_output_open")
        try:
            path = L.enzoe_output_file_path(handle).decode()
            nb = list(root_blocks)
            for i, blk in enumerate(blocks):
                if indices is None:
                    gi = (i % nb[0], (i // nb[0]) % nb[1], i // (nb[0] * nb[1]))
                    lev = 0
                else:
                    gi = indices[i]
                    lev = levels[i] if levels is not None else 0
                n, g = _active(blk, rank)
                arr, nf, keep = _fields(blk, self.field_list)
                ob = OutputBlock(_index(lev, gi), int(cycle), float(time), float(dt),
                                 _ints3(n), _ints3(g), arr, nf, int(bool(blk.on_device)))
                _check(L.enzoe_output_write_block(handle, ctypes.byref(ob), stream),
                       "enzoe_output_write_block")
                del keep
        finally:
            rc = L.enzoe_output_close(handle)
        _check(rc, "enzoe_output_close")
        return path


def read_state(path):
    """(cycle, time, dt) of an output file."""
    c, t, d = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
    _check(io_lib().enzoe_input_read_state(path.encode(), ctypes.byref(c), ctypes.byref(t),
                                           ctypes.byref(d)), "enzoe_input_read_state")
    return c.value, t.value, d.value


def read_block(path, name, block, field_list=None, stream=None):
    """Reads the fields of group /<name> into `block`; returns the block's
    (cycle, time, dt)."""
    arr, nf, keep = _fields(block, field_list)
    c, t, d = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
    _check(io_lib().enzoe_input_read_block(path.encode(), name.encode(), arr, nf,
                                           int(bool(block.on_device)), stream,
                                           ctypes.byref(c), ctypes.byref(t),
                                           ctypes.byref(d)), "enzoe_input_read_block")
    del keep
    return c.value, t.value, d.value


def read_block_list(path):
    """{block name: file name} of a .block_list file."""
    with open(path) as f:
        return dict(line.split() for line in f if line.strip())

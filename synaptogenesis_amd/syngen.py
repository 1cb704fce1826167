"""ctypes host layer for the Synaptogenesis C-ABI (``libsyngen.so``).

Mirrors the reference's Python front end (lib/syngen.py:1-706): the same
class and function names (``Properties``, ``Network``, ``Environment``,
``FloatArray``/``IntArray``, ``create_*_callback``, ``get_cpu``/``get_gpus``)
with the same argument meaning and the same error behaviour (a failed ``run``
prints "Failed to run network!" and yields empty Properties, a missing key
yields a null array).  It is written against the C-ABI declared in
``include/syngen.h`` only, so the same code drives any library that exports
that ABI: the MI355X library built in this package (the default) or the
reference's serial CPU build (pass its path to :func:`load_library`).

Unlike the reference module, nothing is loaded at import time; the library is
opened on first use so that a missing build fails with a clear error.
"""

import ctypes
import os
from collections import OrderedDict
from ctypes import (CFUNCTYPE, POINTER, Structure, c_bool, c_char, c_char_p,
                    c_float, c_int, c_longlong, c_uint, c_void_p, cast,
                    string_at)
from json import dumps

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_LIBRARY = os.path.join(_HERE, "libsyngen.so")

# POINTER_TYPE values of the C-ABI (extern.h:6-12)
FLOAT_POINTER, INT_POINTER, STRING_POINTER, PROPS_POINTER, VOID_POINTER = (
    1, 2, 3, 4, 5)


class CArray(Structure):
    """``struct ARRAY`` returned by value (extern.h:14-19)."""
    _fields_ = [("size", c_int), ("type", c_uint),
                ("data", c_void_p), ("owner", c_bool)]


class PArray:
    """Typed view over memory owned by the library (or by a CArray)."""

    ctype = c_void_p

    def __init__(self, size, ptr):
        self.size = size
        self.data = cast(ptr, POINTER(self.ctype))

    def __getitem__(self, index):
        return self.data[index]

    def __setitem__(self, index, value):
        self.data[index] = value

    def __iter__(self):
        for i in range(self.size):
            yield self.data[i]

    def __len__(self):
        return self.size

    def is_void(self):
        return False

    def to_np_array(self):
        if self.size == 0 or not self.data:
            return np.zeros(0, dtype=np.dtype(self.ctype))
        return np.ctypeslib.as_array(self.data, shape=(self.size,))

    def copy_from(self, arr):
        np.copyto(self.to_np_array(), np.asarray(arr))


class FloatArray(PArray):
    ctype = c_float

    def to_list(self):
        return [float(x) for x in self]


class IntArray(PArray):
    ctype = c_int

    def to_list(self):
        return [int(x) for x in self]


class StringArray(PArray):
    ctype = POINTER(c_char)


class VoidArray(PArray):
    ctype = c_void_p

    def is_void(self):
        return True

    def to_np_array(self):
        raise TypeError("void array has no numeric view")


def build_array(base_array):
    """Wrap a CArray in the typed Python view its ``type`` field names."""
    kind = {FLOAT_POINTER: FloatArray, INT_POINTER: IntArray,
            STRING_POINTER: StringArray, PROPS_POINTER: VoidArray,
            VOID_POINTER: VoidArray}.get(base_array.type, VoidArray)
    return kind(base_array.size, base_array.data)


IO_CALLBACK = CFUNCTYPE(None, c_int, c_int, c_void_p)
WEIGHT_CALLBACK = CFUNCTYPE(None, c_int, c_int, c_void_p)
INDICES_CALLBACK = CFUNCTYPE(None, c_int, c_int, c_void_p, c_void_p,
                             c_void_p, c_void_p)
DISTANCE_CALLBACK = CFUNCTYPE(None, c_int, c_int, c_void_p, c_void_p)
DELAY_CALLBACK = CFUNCTYPE(None, c_int, c_int, c_void_p, c_void_p)


class Library:
    """One opened C-ABI library with its argument and return types set.

    The table follows the binding the reference performs at
    lib/syngen.py:92-175.
    """

    def __init__(self, path):
        if not os.path.exists(path):
            raise OSError(
                "Synaptogenesis library not found at %s; build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'`" % path)
        self.path = path
        self.lib = ctypes.CDLL(path)
        L = self.lib
        sig = {
            "free_array": (None, [CArray]),
            "free_array_deep": (None, [CArray]),
            "create_properties": (c_void_p, []),
            "add_property": (None, [c_void_p, c_char_p, c_char_p]),
            "add_child": (None, [c_void_p, c_char_p, c_void_p]),
            "add_to_array": (None, [c_void_p, c_char_p, c_char_p]),
            "add_to_child_array": (None, [c_void_p, c_char_p, c_void_p]),
            "get_keys": (CArray, [c_void_p]),
            "get_child_keys": (CArray, [c_void_p]),
            "get_array_keys": (CArray, [c_void_p]),
            "get_child_array_keys": (CArray, [c_void_p]),
            "get_property": (c_char_p, [c_void_p, c_char_p]),
            "get_child": (c_void_p, [c_void_p, c_char_p]),
            "get_array": (CArray, [c_void_p, c_char_p]),
            "get_child_array": (CArray, [c_void_p, c_char_p]),
            "create_environment": (c_void_p, [c_void_p]),
            "load_env": (c_void_p, [c_char_p]),
            "save_env": (c_bool, [c_void_p, c_char_p]),
            "create_network": (c_void_p, [c_void_p]),
            "load_net": (c_void_p, [c_char_p]),
            "save_net": (c_bool, [c_void_p, c_char_p]),
            "build_state": (c_void_p, [c_void_p]),
            "build_load_state": (c_void_p, [c_void_p, c_char_p]),
            "load_state": (c_bool, [c_void_p, c_char_p]),
            "save_state": (c_bool, [c_void_p, c_char_p]),
            "get_neuron_data": (CArray, [c_void_p, c_char_p, c_char_p,
                                         c_char_p]),
            "get_layer_data": (CArray, [c_void_p, c_char_p, c_char_p,
                                        c_char_p]),
            "get_connection_data": (CArray, [c_void_p, c_char_p, c_char_p]),
            "get_weight_matrix": (CArray, [c_void_p, c_char_p, c_char_p]),
            "run": (c_void_p, [c_void_p, c_void_p, c_void_p, c_void_p]),
            "destroy": (None, [c_void_p]),
            "destroy_network": (None, [c_void_p]),
            "destroy_environment": (None, [c_void_p]),
            "destroy_state": (None, [c_void_p]),
            "destroy_properties": (None, [c_void_p]),
            "get_cpu": (c_int, []),
            "get_gpus": (CArray, []),
            "get_all_devices": (CArray, []),
            "set_suppress_output": (None, [c_bool]),
            "set_warnings": (None, [c_bool]),
            "set_debug": (None, [c_bool]),
            "interrupt_engine": (None, []),
            "add_io_callback": (None, [c_char_p, c_longlong]),
            "add_weight_callback": (None, [c_char_p, c_longlong]),
            "add_indices_weight_callback": (None, [c_char_p, c_longlong]),
            "add_distance_weight_callback": (None, [c_char_p, c_longlong]),
            "add_delay_weight_callback": (None, [c_char_p, c_longlong]),
            "initialize_mpi": (None, []),
            "get_mpi_rank": (c_int, []),
            "get_mpi_size": (c_int, []),
            "finalize_mpi": (None, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        # Keep ctypes callback objects alive as long as the library is.
        self._callbacks = {}

    def __getattr__(self, name):
        return getattr(self.lib, name)

    # -- callbacks (lib/syngen.py:207-248) ---------------------------------
    def _register(self, kind, adder, name, f):
        cb = kind(f)
        addr = cast(cb, c_void_p).value
        self._callbacks[(adder, name)] = cb
        getattr(self.lib, adder)(name.encode("utf-8"), addr)
        return cb

    def create_io_callback(self, name, f):
        return self._register(IO_CALLBACK, "add_io_callback", name, f)

    def create_weight_callback(self, name, f):
        return self._register(WEIGHT_CALLBACK, "add_weight_callback", name, f)

    def create_indices_weight_callback(self, name, f):
        return self._register(INDICES_CALLBACK,
                              "add_indices_weight_callback", name, f)

    def create_distance_weight_callback(self, name, f):
        return self._register(DISTANCE_CALLBACK,
                              "add_distance_weight_callback", name, f)

    def create_delay_weight_callback(self, name, f):
        return self._register(DELAY_CALLBACK,
                              "add_delay_weight_callback", name, f)


_default = None


def load_library(path=None):
    """Open (or return the already opened) library at ``path``."""
    global _default
    if path is None:
        if _default is None:
            _default = Library(DEFAULT_LIBRARY)
        return _default
    return Library(path)


def _lib(lib):
    return lib if lib is not None else load_library()


def _s(x):
    if isinstance(x, bool):
        return "true" if x else "false"
    return str(x)


class Properties:
    """C-side string tree (PropertyConfig) built from a dict, or read back
    from a handle (lib/syngen.py:252-391)."""

    def __init__(self, props=None, lib=None):
        self.lib = _lib(lib)
        self.properties = OrderedDict()
        self.children = OrderedDict()
        self.arrays = OrderedDict()
        self.child_arrays = OrderedDict()
        self.obj = None
        if props is None:
            props = {}
        if isinstance(props, dict):
            self.obj = self.lib.create_properties()
            for k, v in props.items():
                if v is None:
                    continue
                if isinstance(v, dict):
                    self.add_child(k, v)
                elif isinstance(v, list):
                    for x in v:
                        if isinstance(x, dict):
                            self.add_to_child_array(k, x)
                        else:
                            self.add_to_array(k, _s(x))
                else:
                    self.add_property(k, v)
        elif isinstance(props, int):
            self.obj = props
            self._read_back(props)

    def _keys(self, getter, handle):
        arr_obj = getter(handle)
        arr = build_array(arr_obj)
        keys = [string_at(cast(arr.data[i], c_char_p)).decode("utf-8")
                for i in range(arr.size)]
        self.lib.free_array_deep(arr_obj)
        return keys

    def _read_back(self, h):
        L = self.lib
        for k in self._keys(L.get_keys, h):
            self.properties[k] = L.get_property(h, k.encode()).decode()
        for k in self._keys(L.get_child_keys, h):
            self.children[k] = Properties(L.get_child(h, k.encode()), L)
        for k in self._keys(L.get_array_keys, h):
            arr_obj = L.get_array(h, k.encode())
            arr = build_array(arr_obj)
            self.arrays.setdefault(k, []).extend(
                string_at(cast(arr.data[j], c_char_p)).decode()
                for j in range(arr.size))
            L.free_array_deep(arr_obj)
        for k in self._keys(L.get_child_array_keys, h):
            arr_obj = L.get_child_array(h, k.encode())
            arr = build_array(arr_obj)
            self.child_arrays.setdefault(k, []).extend(
                Properties(arr.data[j], L) for j in range(arr.size))
            L.free_array(arr_obj)

    def to_dict(self):
        out = OrderedDict()
        out.update(self.properties)
        for k, v in self.children.items():
            out[k] = v.to_dict()
        for k, v in self.arrays.items():
            out[k] = [str(p) for p in v]
        for k, v in self.child_arrays.items():
            out[k] = [p.to_dict() for p in v]
        return out

    def __str__(self):
        return dumps(self.to_dict(), indent=4)

    def get(self, key, default=None):
        return self.properties.get(key, default)

    def add_property(self, key, val):
        self.properties[key] = val
        self.lib.add_property(self.obj, key.encode(), _s(val).encode())

    def add_child(self, key, child):
        child = Properties(child, self.lib)
        self.children[key] = child
        self.lib.add_child(self.obj, key.encode(), child.obj)

    def add_to_array(self, key, val):
        self.arrays.setdefault(key, []).append(val)
        self.lib.add_to_array(self.obj, key.encode(), val.encode())

    def add_to_child_array(self, key, dictionary):
        props = Properties(dictionary, self.lib)
        self.child_arrays.setdefault(key, []).append(props)
        self.lib.add_to_child_array(self.obj, key.encode(), props.obj)

    def free(self):
        if self.obj is not None:
            self.lib.destroy_properties(self.obj)
            self.obj = None


class Environment:
    """Environment handle from a JSON path or a dict (lib/syngen.py:395-415)."""

    def __init__(self, env, lib=None):
        self.lib = _lib(lib)
        self.props = None
        if isinstance(env, str):
            self.obj = self.lib.load_env(env.encode())
        elif isinstance(env, dict):
            self.props = Properties(env, self.lib)
            self.obj = self.lib.create_environment(self.props.obj)
        else:
            self.obj = None

    def save(self, filename):
        return self.lib.save_env(self.obj, filename.encode())

    def free(self):
        if self.props is not None:
            self.props.free()
            self.props = None
        if self.obj is not None:
            self.lib.destroy_environment(self.obj)
            self.obj = None


class Network:
    """Network + State handle (lib/syngen.py:419-507)."""

    def __init__(self, net, lib=None):
        self.lib = _lib(lib)
        self.props = None
        if isinstance(net, str):
            self.obj = self.lib.load_net(net.encode())
        elif isinstance(net, dict):
            self.props = Properties(net, self.lib)
            self.obj = self.lib.create_network(self.props.obj)
        else:
            self.obj = None
        self.state = None

    def save(self, filename):
        return self.lib.save_net(self.obj, filename.encode())

    def free(self):
        if self.props is not None:
            self.props.free()
            self.props = None
        if self.state is not None:
            self.lib.destroy_state(self.state)
            self.state = None
        if self.obj is not None:
            self.lib.destroy_network(self.obj)
            self.obj = None

    def build_state(self, filename=None):
        if self.state is not None:
            self.lib.destroy_state(self.state)
        if filename is None:
            self.state = self.lib.build_state(self.obj)
        else:
            self.state = self.lib.build_load_state(self.obj, filename.encode())
        return self.state

    def load_state(self, filename):
        if self.state is None:
            return self.build_state(filename) is not None
        return self.lib.load_state(self.state, filename.encode())

    def save_state(self, filename):
        if self.state is None:
            self.build_state()
        return self.lib.save_state(self.state, filename.encode())

    def get_neuron_data(self, structure, layer, key):
        if self.state is None:
            self.build_state()
        return build_array(self.lib.get_neuron_data(
            self.state, structure.encode(), layer.encode(), key.encode()))

    def get_layer_data(self, structure, layer, key):
        if self.state is None:
            self.build_state()
        return build_array(self.lib.get_layer_data(
            self.state, structure.encode(), layer.encode(), key.encode()))

    def get_connection_data(self, structure, connection, key):
        if self.state is None:
            self.build_state()
        return build_array(self.lib.get_connection_data(
            self.state, connection.encode(), key.encode()))

    def get_weight_matrix(self, connection, key="weights"):
        if self.state is None:
            self.build_state()
        return build_array(self.lib.get_weight_matrix(
            self.state, connection.encode(), key.encode()))

    def run(self, environment=None, args=None):
        if self.state is None:
            self.build_state()
        if not isinstance(environment, Environment):
            environment = Environment(environment, self.lib)
        if not isinstance(args, Properties):
            args = Properties(args or {}, self.lib)
        report = self.lib.run(self.obj, environment.obj, self.state, args.obj)
        if report is None:
            print("Failed to run network!")
            return Properties({}, self.lib)
        return Properties(report, self.lib)


# -- module-level conveniences bound to the default library ----------------

def get_cpu(lib=None):
    return _lib(lib).get_cpu()


def _int_list(arr_obj, lib):
    vals = build_array(arr_obj).to_list() if arr_obj.size else []
    lib.free_array(arr_obj)
    return vals


def get_gpus(lib=None):
    L = _lib(lib)
    return _int_list(L.get_gpus(), L)


def get_all_devices(lib=None):
    L = _lib(lib)
    return _int_list(L.get_all_devices(), L)


def set_suppress_output(val, lib=None):
    _lib(lib).set_suppress_output(bool(val))


def set_warnings(val, lib=None):
    _lib(lib).set_warnings(bool(val))


def set_debug(val, lib=None):
    _lib(lib).set_debug(bool(val))


def interrupt_engine(lib=None):
    _lib(lib).interrupt_engine()


def create_io_callback(name, f, lib=None):
    return _lib(lib).create_io_callback(name, f)


def create_weight_callback(name, f, lib=None):
    return _lib(lib).create_weight_callback(name, f)


def create_indices_weight_callback(name, f, lib=None):
    return _lib(lib).create_indices_weight_callback(name, f)


def create_distance_weight_callback(name, f, lib=None):
    return _lib(lib).create_distance_weight_callback(name, f)


def create_delay_weight_callback(name, f, lib=None):
    return _lib(lib).create_delay_weight_callback(name, f)


def make_custom_input_module(structure, layer_names, name, cb, clear=True,
                             lib=None):
    """Callback input module config (lib/syngen.py:567-586)."""
    def custom_callback(ID, size, ptr):
        cb(layer_names[ID], FloatArray(size, ptr).to_np_array())
    create_io_callback(name, custom_callback, lib)
    return {"type": "callback", "clear": clear,
            "layers": [{"structure": structure, "layer": ln, "input": True,
                        "function": name, "id": i}
                       for i, ln in enumerate(layer_names)]}


def make_custom_output_module(structure, layer_names, name, cb, lib=None):
    """Callback output module config (lib/syngen.py:588-606)."""
    def custom_callback(ID, size, ptr):
        cb(layer_names[ID], FloatArray(size, ptr).to_np_array())
    create_io_callback(name, custom_callback, lib)
    return {"type": "callback",
            "layers": [{"structure": structure, "layer": ln, "output": True,
                        "function": name, "id": i}
                       for i, ln in enumerate(layer_names)]}

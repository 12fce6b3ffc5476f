"""C ABI of libgpuexplore.so on the CPU: every declared symbol is exported,
the gpf loader and model compiler behave, errors are reported as statuses."""

import ctypes
import os
import re

import pytest

from conftest import GOLDEN, ROOT
from gpuexplore_amd import gen

HEADER = os.path.join(ROOT, "include", "gpuexplore.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gx_[a-z0-9_]+)\s*\(", text)))


def test_exports_every_declared_symbol(gx):
    lib = gx.lib()
    names = declared_symbols()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) <= set(gx.SIGNATURES)


def test_load_tiny2_info(gx):
    m = gx.Model.load(os.path.join(GOLDEN, "tiny2"))
    i = m.info()
    assert i["nr_procs"] == 2 and i["bits_act"] == 1 and i["sv_nints"] == 1
    assert i["state_bits"] == 2 and i["key_words"] == 1 and i["nr_trans"] == 4
    m2 = gx.Model.load(os.path.join(GOLDEN, "tiny2"), por=True)
    assert m2.info()["bits_statevector"] == 4  # cu:1912


def test_load_errors(gx, model_dir):
    with pytest.raises(gx.GxError) as e:
        gx.Model.load(os.path.join(model_dir, "missing"))
    assert e.value.status == -1
    txt = open(os.path.join(GOLDEN, "tiny2.gpf")).read()
    with pytest.raises(gx.GxError) as e:  # truncated
        gx.Model.from_text("\n".join(txt.split("\n")[:12]))
    assert e.value.status == -2
    with pytest.raises(gx.GxError) as e:  # 40 processes
        gx.Model.from_text("0\n40\n2\n8\n20\n")
    assert e.value.status == -8


def test_large_values_roundtrip(gx):
    """values up to 2^32-1 are accepted like glibc atoi (cu:1888-1982)"""
    net = gen.Network([2, 2], 16, [[[] for _ in range(2)] for _ in range(2)],
                      [[{} for _ in range(2)] for _ in range(2)], {})
    # bits_act=4 with bs=2 fills entries up to bit 31 when targets are long
    for p in range(2):
        net.local[p][0].append((15, [1]))
        net.local[p][1].append((15, [0]))
    m = gx.Model.from_text(net.to_gpf())
    assert m.info()["bits_act"] == 4


def test_generated_models_compile(gx):
    for net in [gen.config1(1), gen.config2(1), gen.config3(1)]:
        m = gx.Model.from_text(net.to_gpf())
        i = m.info()
        assert i["key_words"] == (i["state_bits"] + 62) // 63


def test_batch_without_device_reports_status(gx):
    if gx.device_count() > 0:
        pytest.skip("a device is visible")
    m = gx.Model.load(os.path.join(GOLDEN, "tiny2"))
    with pytest.raises(gx.GxError) as e:
        gx.explore_batch([m, m], [0, 1])
    assert e.value.status == -6


def test_explore_without_device_reports_status(gx):
    if gx.device_count() > 0:
        pytest.skip("a device is visible")
    m = gx.Model.load(os.path.join(GOLDEN, "tiny2"))
    with pytest.raises(gx.GxError) as e:
        gx.explore(m)
    assert e.value.status == -6


def test_server_without_device_reports_status(gx):
    if gx.device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(gx.GxError) as e:
        gx.Server(slots=4)
    assert e.value.status == -6


def test_server_options_default(gx):
    o = gx.ServerOptions()
    gx.lib().gx_server_options_default(o)
    assert o.slots == 256 and o.key_words == 2 and o.rounds_per_step >= 1
    assert o.table_bytes % o.page_bytes == 0


@pytest.mark.parametrize("cls,fn", [("Options", "gx_options_default"),
                                    ("BatchOptions", "gx_batch_options_default"),
                                    ("ServerOptions", "gx_server_options_default")])
def test_option_struct_layouts(gx, cls, fn):
    """the ctypes mirrors have the C layouts: the C defaults function writes
    exactly the mirror's bytes (guard bytes after it untouched) and every
    field, the last one included (the struct was filled with 0xFF first)"""
    S = getattr(gx, cls)

    class Guarded(ctypes.Structure):
        _fields_ = [("o", S), ("guard", ctypes.c_uint8 * 64)]

    g = Guarded()
    ctypes.memset(ctypes.addressof(g), 0xFF, ctypes.sizeof(S))
    ctypes.memset(ctypes.addressof(g.guard), 0xA5, 64)
    getattr(gx.lib(), fn)(ctypes.byref(g.o))
    assert list(g.guard) == [0xA5] * 64
    last = S._fields_[-1][0]
    assert getattr(g.o, last) != (1 << (8 * ctypes.sizeof(S._fields_[-1][1]))) - 1, last
    if cls == "Options":
        assert g.o.kernel_iters == 1 and g.o.por_method == gx.POR_CAMPLE and g.o.hash_seed != 0


def test_unknown_option_names_are_errors(gx):
    """a misspelled option must not be silently ignored"""
    assert gx.options(por_method=gx.POR_STUBBORN, q_ints=1 << 20).por_method == gx.POR_STUBBORN
    with pytest.raises(TypeError):
        gx.options(por_methd=2)
    with pytest.raises(TypeError):
        gx.explore_batch([], [], not_a_field=1)
    with pytest.raises(TypeError):
        gx.Server(slot=4)


def test_null_pointers_are_statuses(gx):
    """null handles and out-pointers give GX_E_ARG (or nothing, for the
    void defaults functions) instead of crashing the caller"""
    L = gx.lib()
    L.gx_options_default(None)
    L.gx_batch_options_default(None)
    L.gx_server_options_default(None)
    assert L.gx_model_load_gpf(None, 0, None) == -7
    assert L.gx_model_info_get(None, None) == -7
    assert L.gx_explorer_run(None, None) == -7
    assert L.gx_explorer_dump(None, None) == -7
    assert L.gx_device_count(None) == -7

"""Every C-ABI entry point answers invalid arguments with a negative status
and a message from hmg_last_error, without a GPU and without touching
memory: null contexts, sessions, batches and servers, null output
pointers, zero, negative and unknown sizes and codes.  The only calls that
succeed are the documented no-ops: destroying a null handle, a remap of
zero elements, and selecting kernel choice 0."""
import ctypes as C

import pytest

OK_NOOPS = {("hmg_destroy", "*"), ("hmg_session_destroy", "*"), ("hmg_batch_destroy", "*"),
            ("hmg_server_destroy", "*"), ("hmg_remap_dev", "zero"), ("hmg_remap_dev", "neg"),
            ("hmg_thermal_set_block_steps", "zero")}
NOT_STATUS = {"hmg_last_error", "hmg_abi_version", "hmg_vpipes_defaults", "hmg_stream",
              "hmg_session_z_dev", "hmg_session_bytes"}


def _entry_points():
    import highmap_amd
    from highmap_amd import serve

    L = serve._bind()
    names = sorted(n for n in dir(L) if n.startswith("hmg_") and not n.startswith("_"))
    if not names:  # ctypes binds lazily: take the names the wrappers declared
        import re

        src = open(highmap_amd.__file__).read() + open(serve.__file__).read()
        names = sorted(set(re.findall(r"\b(hmg_[a-z_]+)\b", src)))
    return L, [n for n in names if n not in NOT_STATUS and getattr(L, n).argtypes is not None]


def _args(fn, variant):
    out = []
    for t in fn.argtypes:
        if t in (C.c_int, C.c_uint32, C.c_longlong):
            out.append({"zero": 0, "shape": 64, "neg": -5}[variant])
        elif t in (C.c_float, C.c_double):
            out.append(0.5)
        elif t is C.c_size_t:
            out.append({"zero": 0, "shape": 64, "neg": 0}[variant])
        else:
            out.append(None)
    return out


@pytest.mark.parametrize("variant", ["zero", "shape", "neg"])
def test_invalid_arguments_are_errors(variant):
    L, names = _entry_points()
    assert len(names) >= 40, names
    try:
        for n in names:
            fn = getattr(L, n)
            rc = fn(*_args(fn, variant))
            if (n, "*") in OK_NOOPS or (n, variant) in OK_NOOPS:
                assert rc == 0, (n, variant, rc)
                continue
            assert rc < 0, (n, variant, rc)
            assert L.hmg_last_error(), (n, variant)
    finally:
        assert L.hmg_thermal_set_block_steps(8) == 0

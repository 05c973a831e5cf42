"""CPU tests of the ork texture descriptors the renderers load the tables
through (atmo_texture_load_desc / atmo_texture_load, include/atmo_consumer.h).

Descriptor form: src/demo/textures/atmo/inscatter.xml:25 and
src/atmo/examples/atmo/helloworld.xml:71-75 (a texture2D / texture3D element
with source, internalformat, min, mag and wrap attributes, after comment and
<?xml ?> lines).  The RGBA16F payload is checked bit for bit against NumPy's
IEEE binary16 conversion (round to nearest even), an implementation
independent of the library's.
"""
import os

import numpy as np
import pytest

from prolandatmo import consumer as C
from prolandatmo import lib as L

HEAD = '<!--\n * a comment header, as the reference files carry\n-->\n<?xml version="1.0" ?><?xml version="1.0" ?>\n'


def write_desc(d, name, kind, source, fmt="RGBA16F", wraps=("CLAMP_TO_EDGE",) * 3, minf="LINEAR"):
    axes = ("wraps", "wrapt", "wrapr")[: (3 if kind == "texture3D" else 2)]
    w = " ".join('%s="%s"' % (a, m) for a, m in zip(axes, wraps) if m is not None)
    path = os.path.join(d, name + ".xml")
    with open(path, "w") as f:
        f.write(HEAD + '<%s name="%s" source="%s" internalformat="%s" min="%s" mag="LINEAR" %s/>\n'
                % (kind, name, source, fmt, minf, w))
    return path


def tables(tmp_path):
    rng = np.random.default_rng(7)
    T = rng.random((16, 64, 3)).astype(np.float32)
    S = (rng.random((4, 16, 8 * 4, 4)) * np.float32(3e-3)).astype(np.float32)
    d = str(tmp_path)
    L.write_raw(os.path.join(d, "transmittance.raw"), T, 64, 16, 0, 3)
    L.write_raw(os.path.join(d, "inscatter.raw"), S, 32, 16, 4, 4)
    return d, T, S


def test_descriptors_of_the_reference_form(tmp_path):
    d, T, S = tables(tmp_path)
    x2 = write_desc(d, "transmittance", "texture2D", "transmittance.raw")
    x3 = write_desc(d, "inscatter", "texture3D", "inscatter.raw")
    t = C.texture_desc(x2)
    assert (t["kind"], t["width"], t["height"], t["depth"], t["components"]) == (2, 64, 16, 0, 3)
    assert t["format"] == "RGBA16F" and t["half"] == 1 and t["format_channels"] == 4
    assert t["min_linear"] and t["mag_linear"] and t["clamp_to_edge"][:2] == [1, 1]
    assert t["source"] == os.path.join(d, "transmittance.raw") and t["name"] == "transmittance"
    s = C.texture_desc(x3)
    assert (s["kind"], s["width"], s["height"], s["depth"], s["components"]) == (3, 32, 16, 4, 4)
    assert s["clamp_to_edge"] == [1, 1, 1]
    a, _ = C.load_texture(x2)
    assert a.shape == (16, 64, 3)
    assert np.array_equal(a.view(np.uint32), T.astype(np.float16).astype(np.float32).view(np.uint32))
    b, _ = C.load_texture(x3)
    assert b.shape == (4, 16, 32, 4)
    assert np.array_equal(b.view(np.uint32), S.astype(np.float16).astype(np.float32).view(np.uint32))
    # a 32F format keeps the values
    x4 = write_desc(d, "t32", "texture2D", "transmittance.raw", fmt="RGB32F")
    c, _ = C.load_texture(x4)
    assert np.array_equal(c, T)


def test_half_rounding_matches_ieee_binary16(tmp_path):
    rng = np.random.default_rng(3)
    v = np.concatenate([
        rng.standard_normal(20000) * np.exp(rng.uniform(-25, 12, 20000)),   # subnormal .. overflow
        np.array([0.0, -0.0, 65504.0, 65519.99, 65520.0, -65520.0, 6.1035156e-05, 6.09e-05,
                  2.9802322e-08, 2.9802326e-08, 5.960464e-08, 1.0 + 2 ** -11, 1.0 + 3 * 2 ** -11,
                  np.inf, -np.inf])]).astype(np.float32)
    d = str(tmp_path)
    L.write_raw(os.path.join(d, "v.raw"), v, v.size, 1, 0, 1)
    x = write_desc(d, "v", "texture2D", "v.raw", fmt="R16F")
    a, _ = C.load_texture(x)
    with np.errstate(over="ignore"):
        want = v.astype(np.float16).astype(np.float32)
    assert np.array_equal(a.ravel().view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("case", ["kind2d_on_3d", "kind3d_on_2d", "format", "channels", "no_wrap",
                                  "missing_source", "bad_trailer", "no_element"])
def test_descriptors_that_do_not_match_are_rejected(tmp_path, case):
    d, T, S = tables(tmp_path)
    if case == "kind2d_on_3d":
        x, st = write_desc(d, "a", "texture2D", "inscatter.raw"), 4
    elif case == "kind3d_on_2d":
        x, st = write_desc(d, "a", "texture3D", "transmittance.raw"), 4
    elif case == "format":
        x, st = write_desc(d, "a", "texture2D", "transmittance.raw", fmt="RGB8"), 6
    elif case == "channels":
        x, st = write_desc(d, "a", "texture3D", "inscatter.raw", fmt="RGB16F"), 6
    elif case == "no_wrap":
        x, st = write_desc(d, "a", "texture2D", "transmittance.raw", wraps=("CLAMP_TO_EDGE", None)), 6
    elif case == "missing_source":
        x, st = write_desc(d, "a", "texture2D", "nothere.raw"), 4
    elif case == "bad_trailer":
        p = os.path.join(d, "cut.raw")
        with open(os.path.join(d, "transmittance.raw"), "rb") as f:
            data = f.read()
        with open(p, "wb") as f:
            f.write(data[:-24] + data[-20:])   # payload one texel short
        x, st = write_desc(d, "a", "texture2D", "cut.raw"), 4
    else:
        x = os.path.join(d, "a.xml")
        with open(x, "w") as f:
            f.write(HEAD + '<module name="m"/>\n')
        st = 4
    with pytest.raises(L.AtmoError) as e:
        C.texture_desc(x)
    assert e.value.status == st, str(e.value)


def test_sampler_state_is_reported(tmp_path):
    d, T, S = tables(tmp_path)
    x = write_desc(d, "a", "texture2D", "transmittance.raw", wraps=("REPEAT", "CLAMP_TO_EDGE"),
                   minf="NEAREST")
    t = C.texture_desc(x)
    assert t["clamp_to_edge"][:2] == [0, 1] and t["min_linear"] == 0 and t["mag_linear"] == 1

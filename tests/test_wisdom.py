"""Wisdom cache of the autotuner, host side only (no device calls):
import of a wisdom file, export of the cache, forgetting, and rejection of
malformed files."""
import pytest

import pfft_amd as P

try:
    P.lib()
    HAVE_LIB = True
except Exception:   # library not built
    HAVE_LIB = False

pytestmark = pytest.mark.skipif(not HAVE_LIB, reason="libpfft.so not built")

GOOD = """pfft-amd-wisdom 1 2
k0.p0.n512.f512.i512@0.o512@0.s1:1.t512:512:512.h0.m0.mo00.b512:262144:262144 2 0 1.250000
k1.p0.n512.f256.i512@0.o257@0.s1:1.t512:512:264.h0.m0.mo00.b512:262144:135168 0 1 0.900000
"""


def test_import_export_roundtrip(tmp_path):
    P.forget_wisdom()
    assert P.wisdom_count() == 0
    f = tmp_path / "w.txt"
    f.write_text(GOOD)
    assert P.import_wisdom(str(f)) == 0
    assert P.wisdom_count() == 2
    g = tmp_path / "w2.txt"
    assert P.export_wisdom(str(g)) == 0
    assert g.read_text() == GOOD
    # importing the same selections again replaces, does not duplicate
    assert P.import_wisdom(str(g)) == 0
    assert P.wisdom_count() == 2
    P.forget_wisdom()
    assert P.wisdom_count() == 0


@pytest.mark.parametrize("text", [
    "not-wisdom 1 1\nk 0 0 1.0\n",
    "pfft-amd-wisdom 1 2\nkey 0 0 1.0\n",             # fewer lines than announced
    "pfft-amd-wisdom 1 1\nkey 5 0 1.0\n",             # unknown variant
    "pfft-amd-wisdom 1 1\nkey 0 2 1.0\n",             # remap is 0 or 1
    "pfft-amd-wisdom 1 1\nkey zero 0 1.0\n",
])
def test_malformed_rejected(tmp_path, text):
    P.forget_wisdom()
    f = tmp_path / "bad.txt"
    f.write_text(text)
    assert P.import_wisdom(str(f)) == 1


def test_missing_file(tmp_path):
    assert P.import_wisdom(str(tmp_path / "none.txt")) == 1
    assert P.export_wisdom(str(tmp_path / "no_dir" / "w.txt")) == 1

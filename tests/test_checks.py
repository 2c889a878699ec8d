"""The reference's behavioural suite (52 checks, test/test_main.cpp) on both
engines: the reference build pins the restatement, the device run is the
parity contract."""
import pytest

import m3d_checks as K

NAMES = [f.__name__ for f in K.ALL]


def _run(lib, name):
    fn = getattr(K, name)
    failed = [msg for msg, ok in fn(lib) if not ok]
    assert not failed, f"{name}: " + "; ".join(failed)


def test_check_count_matches_reference(ref):
    assert sum(len(f(ref)) for f in K.ALL) == 52


@pytest.mark.parametrize("name", NAMES)
def test_reference_checks(ref, name):
    _run(ref, name)


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_device_checks(dev, name):
    _run(dev, name)

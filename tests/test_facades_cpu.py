"""Facade argument handling that fails before any device work (CPU only)."""
import pytest

import padeops_amd as po


@pytest.mark.parametrize("methods,code", [(("cd10", "four", "cd06"), 13), (("cheb", "cd10", "cd10"), 14),
                                          (("cd10", "cd10", "cd8"), 1)])
def test_derivative_methods(methods, code):
    with pytest.raises(po.PadeError) as e:
        po.derivatives().init(16, 16, 16, 0.1, 0.1, 0.1, True, True, True, *methods)
    assert e.value.code == code


@pytest.mark.parametrize("kw,code", [({"xmetric": True}, 4), ({"ymetric": True}, 5), ({"zmetric": True}, 6),
                                     ({"curvilinear": True}, 6)])
def test_metric_flags(kw, code):
    with pytest.raises(po.PadeError) as e:
        po.derivatives().init(16, 16, 16, 0.1, 0.1, 0.1, True, True, True, "cd10", "cd10", "cd10", **kw)
    assert e.value.code == code


@pytest.mark.parametrize("methods,per", [(("gaussian", "lstsq", "cf90"), (False, True, True)),
                                         (("cf90", "lstsq", "cf90"), (True, True, True)),
                                         (("cf90", "cf90", "box"), (True, True, True))])
def test_filter_methods(methods, per):
    # the Gaussian filter is provided; the least-squares filter and unknown
    # names stop with 52 (filters.F90:120)
    with pytest.raises(po.PadeError) as e:
        po.filters().init(16, 16, 16, *per, *methods)
    assert e.value.code == 52


def test_gp_record_sizes():
    from padeops_amd.facades import _sizes

    class GP:
        xsz = (8, 10, 12)
        ysz = (8, 10, 12)
        zsz = (8, 10, 12)

    assert _sizes((GP(), 1, 2))[:3] == (8, 10, 12)
    assert _sizes((8, 9, 10, 1))[:3] == (8, 9, 10)

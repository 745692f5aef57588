"""CPU restatement of ODTbrain's object-function translations (TEST
INFRASTRUCTURE: a checker, never imported by the product package).

odt_to_ri restates odtbrain/_translate_ri.py:5-57 and opt_to_ri
_translate_ri.py:60-90 in numpy.  Pinned by the reference's fixtures
processing__test_odt_to_ri / processing__test_opt_to_ri
(tests/test_processing.py:52-76) and by the inversion identity
f = km^2 ((n / nm)^2 - 1) -> n (tests/test_oracle.py).
"""
import numpy as np


def odt_to_ri(f, res, nm):
    """n = nm sqrt(f / km^2 + 1), km = 2 pi nm / res; roots with negative
    real part negated (_translate_ri.py:47-56)."""
    km = (2 * np.pi * nm) / res
    ri = nm * np.sqrt(f / km ** 2 + 1)
    neg = ri.real < 0
    ri[neg] *= -1
    return ri


def opt_to_ri(f, res, nm):
    """n = nm + f / (2 pi) * res (_translate_ri.py:89)."""
    return nm + f / (2 * np.pi) * res


def processing_test_data():
    """The input of tests/test_processing.py:14-21: a 10^3 ramp plus a
    linspace along the last axis, res = 7, nm = 1.34."""
    ln = 10
    f = np.arange(ln ** 3).reshape(ln, ln, ln)
    f = f + np.linspace(1, 2, ln)
    return f, 7, 1.34

"""Non-parametric CDF data, a Python 3 drop-in for the reference's ``geo_bsd/cdf.py``."""
import numpy


class CdfData:
    def __init__(self, values, probs):
        self.values = numpy.require(values, "float32")
        self.probs = numpy.require(probs, "float32")


def calc_cdf(prop):
    """Cumulative frequencies of the distinct informed values (cdf.py:9-36).

    As in the reference, the largest value is dropped from the table when
    there is more than one distinct value.
    """
    data = numpy.asarray(prop.data)
    mask = numpy.asarray(prop.mask)
    vals = data[mask != 0]
    full_count = float(vals.size)
    values, counts = numpy.unique(vals, return_counts=True)
    if values.size == 0:
        size = 0
    elif values.size == 1:
        size = 1
    else:
        size = values.size - 1
    values = values[:size]
    probs = numpy.zeros(values.shape)
    last_prob = 0.0
    for i in range(size):
        probs[i] = last_prob + counts[i] / full_count
        last_prob = probs[i]
    return CdfData(values=values, probs=probs)

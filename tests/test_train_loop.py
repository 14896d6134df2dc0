"""Host parts of the training loop (CPU only): the batch loss of the validation monitor, the
progress line and the iteration trajectory file.

* batch_loss against closed forms of loss.F90:370-700 on a hand-made set: two structures
  (1 and 2 atoms), one global and one atomic output, datapoint and atomic weights;
* the progress line against the Fortran edit descriptors (I10,5X,E15.6,4X,E15.6,4X,E15.6)
  of bpnn.F90:329-338, and iterout.dat against (I<w>,3ES26.16E3) of iterout.F90.
"""
import numpy as np
import pytest

from fortnet_amd.train import batch_loss, progress_line, write_iterout

OFF = np.array([0, 1, 3])
ES = np.array([[1.0, 0.0], [4.0, 0.0]])          # global output (column 0) per structure
EA = np.array([[1.0, 2.0], [1.5, 3.0], [2.5, -1.0]])  # atomic output in column 1
GT = np.array([[2.0], [1.0]])                     # global targets
AT = np.array([[1.0], [1.0], [1.0]])              # atomic targets
AW = np.array([1.0, 2.0, 0.5])
DW = np.array([3.0, 1.0])


def test_batch_loss_closed_forms():
    # global residuals t - p: 1, -3; atomic: -1, -2, 2
    # mse: (w_s sum aw_s (t - p)^2 + sum w_s aw_i (t - p)^2) / (sum w_s sum aw_s + sum w_s aw_i)
    num = 3 * 1 * 1 + 1 * 2.5 * 9 + (3 * 1 * 1 + 1 * 2 * 4 + 1 * 0.5 * 4)
    den = 3 * 1 + 1 * 2.5 + (3 * 1 + 1 * 2 + 1 * 0.5)
    assert abs(batch_loss(ES, EA, OFF, 1, GT, AT, AW, DW, "mse") - num / den) < 1e-14
    num = 3 * 1 * 1 + 1 * 2.5 * 3 + (3 * 1 * 1 + 1 * 2 * 2 + 1 * 0.5 * 2)
    assert abs(batch_loss(ES, EA, OFF, 1, GT, AT, AW, DW, "mae") - num / den) < 1e-14
    assert abs(batch_loss(ES, EA, OFF, 1, GT, AT, AW, DW, "rms") - num / den) < 1e-14
    num = 100 * (3 * 1 * 0.5 + 1 * 2.5 * 3 + (3 * 1 * 1 + 1 * 2 * 2 + 1 * 0.5 * 2))
    assert abs(batch_loss(ES, EA, OFF, 1, GT, AT, AW, DW, "mape") - num / den) < 1e-12
    # no weights: plain means over items
    assert abs(batch_loss(ES, EA, OFF, 0, None, np.c_[GT[[0, 1, 1]], AT], None, None, "mse")
               - np.mean(np.sum((np.c_[GT[[0, 1, 1]], AT] - EA) ** 2, axis=1) / 2)) < 1e-14
    # global targets only: each structure's loss weighted by its summed atomic weights (atoms)
    ls = np.sqrt(np.sum((np.c_[GT, GT] - ES) ** 2, axis=1) / 2)
    assert abs(batch_loss(ES, EA, OFF, 2, np.c_[GT, GT], None, None, None, "rms")
               - (1 * ls[0] + 2 * ls[1]) / 3) < 1e-14


def test_progress_line_format():
    line = progress_line(12, 1.2345678e-3, 45.0, 0.0)
    assert line == "        12" + " " * 5 + "   0.123457E-02" + " " * 4 + "   0.450000E+02" + \
        " " * 4 + "   0.000000E+00"
    assert progress_line(1, -0.5, 1.0) == "         1" + " " * 5 + "  -0.500000E+00" + \
        " " * 4 + "   0.100000E+01"


def test_iterout_format(tmp_path):
    p = tmp_path / "iterout.dat"
    write_iterout(p, [1.0] * 10, [0.25] * 10, [-3.0e-120] * 10)
    lines = open(p).read().splitlines()
    assert len(lines) == 10
    assert lines[0] == " 1" + "   1.0000000000000000E+000" + "   2.5000000000000000E-001" + \
        "  -3.0000000000000001E-120"
    assert lines[9].startswith("10   1.0000000000000000E+000")
    with pytest.raises(ValueError):
        write_iterout(p)

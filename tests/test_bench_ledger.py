"""Host logic of bench.py's traffic ledger: the counter rows of a rocprofv3
--pmc pass are split into the last graph launch and the calibration copy."""
import pytest

import bench


def rows(seq):
    """rocprofv3 counter_collection rows: (dispatch id, kernel, KiB)."""
    return [{"Dispatch_Id": str(i), "Kernel_Name": n, "Counter_Name": "FETCH_SIZE", "Counter_Value": str(v)}
            for i, n, v in seq]


STEP = ["void shx::k_a<512>(float*)", "void shx::k_b<512, 4>(float*)", "k_c(float*)"]


def launch(first, scale=1.0):
    return [(first + i, n, 100.0 * (i + 1) * scale) for i, n in enumerate(STEP)]


def test_last_launch_and_copy_are_separated():
    seq = [(1, "at::native::fill", 5.0), (2, "void shx::k_setup(float*)", 7.0)]
    seq += launch(10) + launch(20) + launch(30, 2.0) + [(40, "__amd_rocclr_copyBuffer", 128.0)]
    disp = bench._dispatches(rows(seq))
    step, copy = bench._split_launch(disp, len(STEP))
    assert [n for _, n, _ in step] == STEP
    assert [v for _, _, v in step] == [200.0 * 1024, 400.0 * 1024, 600.0 * 1024]
    assert [v for _, _, v in copy] == [128.0 * 1024]


def test_rows_of_one_dispatch_are_summed():
    disp = bench._dispatches(rows([(3, "k_x", 1.0), (3, "k_x", 2.5)]))
    assert disp == [(3, "k_x", 3.5 * 1024)]


def test_mismatched_launches_are_rejected():
    seq = launch(10) + [(13, "void shx::k_other(float*)", 1.0)] + launch(20)[1:] + [(40, "copyBuffer", 1.0)]
    with pytest.raises(RuntimeError):
        bench._split_launch(bench._dispatches(rows(seq)), len(STEP))


def test_missing_copy_is_rejected():
    with pytest.raises(RuntimeError):
        bench._split_launch(bench._dispatches(rows(launch(10) + launch(20))), len(STEP))

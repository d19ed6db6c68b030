"""Resource admission (admission.rs:17-330) through the C ABI, host only.

The reference's own cases (admission.rs:347-390): a dense request of 10 DOFs,
2 right-hand sides and 3 frequencies has a 10*10*16-byte operator and 6 work
items and fits a budget of exactly its estimate; an over-budget request and an
overflowing one are typed errors; zero work and a zero host budget are
rejected.  Plus the estimate's parts as the reference documents them.
"""

import pytest

import sonium_amd as sa


def test_dense_estimate_and_exact_budget():  # admission.rs:352-361
    e = sa.admission_estimate(10, 2, 3)
    assert e["operator_bytes"] == 10 * 10 * 16
    assert e["active_work"] == 6
    # the documented parts: operator, right-hand sides per frequency, 4 complex + 1 pointer of scratch per DOF
    assert e["rhs_bytes"] == 10 * 2 * 16 * 3 and e["scratch_bytes"] == 10 * (4 * 16 + 8)
    assert e["host_bytes"] == e["operator_bytes"] + e["rhs_bytes"] + e["scratch_bytes"]
    ok = sa.admission_admit(e["host_bytes"], e["device_bytes"], 6, 10, 2, 3)
    assert ok == e


def test_over_budget_and_overflow_are_typed():  # admission.rs:363-376
    with pytest.raises(sa.SbhError) as exc:
        sa.admission_admit(1, 0, 1, 100, 1, 1)
    assert exc.value.status == 4 and "host memory" in str(exc.value)
    e = sa.admission_estimate(100)
    with pytest.raises(sa.SbhError) as exc:
        sa.admission_admit(e["host_bytes"], e["device_bytes"] - 1, 1, 100)
    assert exc.value.status == 4 and "device memory" in str(exc.value)
    e2 = sa.admission_estimate(100, 2, 3)
    with pytest.raises(sa.SbhError) as exc:
        sa.admission_admit(e2["host_bytes"], e2["device_bytes"], 5, 100, 2, 3)
    assert exc.value.status == 4 and "active work" in str(exc.value)
    with pytest.raises(sa.SbhError) as exc:
        sa.admission_estimate(2 ** 62, 2, 1)
    assert exc.value.status == 1 and "overflow" in str(exc.value)


def test_zero_work_and_zero_budget_do_not_pass():  # admission.rs:378-390
    with pytest.raises(sa.SbhError) as exc:
        sa.admission_admit(1, 0, 1, 0, 1, 1)
    assert exc.value.status == 1 and "active_dofs" in str(exc.value)
    with pytest.raises(sa.SbhError) as exc:
        sa.admission_admit(0, 0, 1, 10, 1, 1)
    assert exc.value.status == 1 and "max_host_bytes" in str(exc.value)


def test_device_estimate_covers_the_dense_path():
    # system (matrix and one right-hand side), blocks, GMRES(100) basis
    n = 20480
    e = sa.admission_estimate(n, 1, 1, 100)
    assert e["device_bytes"] >= 16 * n * n + 16 * n * 101
    e64 = sa.admission_estimate(40960, 64, 1, 100)
    assert e64["device_bytes"] >= 16 * 40960 ** 2 + 16 * 40960 * 102 * 64

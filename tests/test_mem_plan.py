"""Device-memory plans (fc_op_device_bytes): the per-op byte formula the
host drivers are checked against on the GPU (tests/test_gpu_mem_plan.py) and
that gates each call on free HBM. These CPU checks pin the argument contract
and properties that hold by construction of the drivers' buffers, without
restating the formula: the reference's calibration pattern is
benchmarks/results/poc3_calibration_v1.md:8-13."""
import capi


def g(b):
    """256-byte granule of one allocation."""
    return max(256, (b + 255) // 256 * 256)


def test_contract_codes():
    ob = capi.op_bytes
    assert ob(capi.OP_CS_RANK, 0, 5) == -1
    assert ob(capi.OP_CS_RANK, 1, (1 << 24) + 1) == -2
    assert ob(capi.OP_CS_RANK, 1 << 16, 1 << 16) == -3
    assert ob(capi.OP_ROLLING_IC, 3, 5, vb=2) == -1
    assert ob(capi.OP_STOCK_CORR, 3, 50000) == -3
    assert ob(capi.OP_STOCK_CORR, 3, 5, precision=7) == -1
    assert ob(capi.OP_FACTOR_CORR, 3, 5, F=129) == -3
    assert ob(99, 3, 5) == -1


def test_granules_and_optional_buffers():
    T, N = 37, 1001
    cells = T * N
    base = capi.op_bytes(capi.OP_CS_RANK, T, N)
    assert base > 0 and base % 256 == 0
    # a mask adds exactly one byte plane
    assert capi.op_bytes(capi.OP_CS_RANK, T, N, flags=capi.PLAN_MASK) - base == g(cells)
    # parameter_scan: each extra unmasked group adds one float plane
    one = capi.op_bytes(capi.OP_PARAMETER_SCAN, T, N, flags=1 << 9)   # group 1 only
    two = capi.op_bytes(capi.OP_PARAMETER_SCAN, T, N, flags=(2 | 8) << 8)
    assert two - one == g(4 * cells)
    # all groups (0 = default) include the mask plane of the masked groups
    assert capi.op_bytes(capi.OP_PARAMETER_SCAN, T, N) == \
        capi.op_bytes(capi.OP_PARAMETER_SCAN, T, N, flags=15 << 8)
    # rolling_ic: rank planes are two f64 planes; f64 values double the inputs
    ic = capi.op_bytes(capi.OP_ROLLING_IC, T, N)
    assert capi.op_bytes(capi.OP_ROLLING_IC, T, N, flags=capi.PLAN_RANKS) - ic == 2 * g(8 * cells)
    assert capi.op_bytes(capi.OP_ROLLING_IC, T, N, vb=8) - ic == 2 * (g(8 * cells) - g(4 * cells))


def test_correlation_plans_grow_with_the_output():
    T = 64
    a = capi.op_bytes(capi.OP_STOCK_CORR, T, 500)
    b = capi.op_bytes(capi.OP_STOCK_CORR, T, 1000)
    assert b - a >= 8 * (1000 * 1000 - 500 * 500)   # the (N,N) f64 output alone
    f12 = capi.op_bytes(capi.OP_FACTOR_CORR, T, 300, F=12)
    f13 = capi.op_bytes(capi.OP_FACTOR_CORR, T, 300, F=13)
    assert f13 > f12
    # the masked variant adds one byte per row
    assert capi.op_bytes(capi.OP_FACTOR_CORR, T, 300, F=12, flags=capi.PLAN_MASK) - f12 == g(T * 300)

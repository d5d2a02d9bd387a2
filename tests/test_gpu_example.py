"""The reference's default 384^3 example on the GPU, 1000 steps: the
kernel__vmax trace must equal the FP32 oracle's bit patterns at every report
step (fixture tests/golden/example_vmax_oracle.txt), and so reproduce the
trace the reference printed in example/example.out."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from example_case import NT, es92, example_config, report_steps
from gpu_helpers import gpu_from_model
from test_oracle_example import oracle_trace, reference_trace

pytestmark = pytest.mark.gpu


def test_example_vmax_trace_bit_exact(orc):
    fix, ref = oracle_trace(32), reference_trace()
    m = O.build_model(example_config())
    g = gpu_from_model(m)
    it = 1
    for rs in report_steps(NT):
        if rs > it:
            g.step(it, rs - it, use_graph=True, block=rs - it)
            it = rs
        v = g.kernel_vmax()
        assert np.array_equal(v, fix[rs][0]), (rs, v, fix[rs][0])
        va = v * np.float32(m.UC) * np.float32(m.M0)
        assert tuple(es92(x) for x in va) == ref[rs], rs
    g.close()

"""The GPU filter's expression compiler and program evaluator, run on the
host (vgpu_eval_host), against the oracle: same predicates as the GPU
filter parity test, selection vectors bit-exact.  No GPU needed."""
import numpy as np
import pytest

from test_gpu_filter import PREDS, STR_PREDS, make_str_table, make_table


@pytest.mark.parametrize("pi", range(len(PREDS)))
def test_compiled_predicate_matches_oracle(native, oracle, pi):
    mk = make_table(native, 5003, 100 + pi, 997)
    got = native.host_eval_batches(mk(), PREDS[pi])
    exp = [b["sel"] for b in native.next_batches(native.filter_node("oracle", mk().scan(), PREDS[pi]))]
    assert len(got) == len(exp)
    for g, e in zip(got, exp):
        assert np.array_equal(g, e)


@pytest.mark.parametrize("pi", range(len(STR_PREDS)))
def test_compiled_string_predicate_matches_oracle(native, oracle, pi):
    """string comparisons compiled to a -1 / 0 / 1 column (computed on the
    host here) compared with 0 by the program"""
    mk = make_str_table(native, 3001, 500 + pi, 700)
    got = native.host_eval_batches(mk(), STR_PREDS[pi])
    exp = [b["sel"] for b in native.next_batches(native.filter_node("oracle", mk().scan(), STR_PREDS[pi]))]
    assert len(got) == len(exp)
    for g, e in zip(got, exp):
        assert np.array_equal(g, e)


def test_compile_errors(native):
    t = native.Table({"x": np.array([1., 2.]), "b": np.array([True, False])})
    with pytest.raises(native.VectraError, match="column not found: zz"):
        native.host_eval_batches(t, ("cmp", ">", ("col", "zz"), ("dbl", 1)))
    with pytest.raises(native.VectraError, match="bool: operands must be boolean"):
        native.host_eval_batches(t, ("and", ("col", "x"), ("lgl", True)))
    with pytest.raises(native.VectraError, match="bool not: operand must be boolean"):
        native.host_eval_batches(t, ("not", ("col", "x")))
    with pytest.raises(native.VectraError, match="negate: unsupported type 2"):
        native.host_eval_batches(t, ("cmp", ">", ("neg", ("col", "b")), ("dbl", 0)))
    with pytest.raises(native.VectraError, match="filter predicate must evaluate to boolean"):
        native.host_eval_batches(t, ("col", "x"))

"""The TORCH_LIBRARY(xint) operator library (xpg_amd/csrc/xint_ops.cpp): it loads next to the kernel library,
registers the five operators with their schemas, and refuses host tensors before any launch."""
import pytest
import torch

SCHEMAS = {
    "extract": "xint::extract(Tensor frames, int frame0, int min_valid, int flag_value, int overload, "
               "bool use_flag_value, Tensor?[] rows, int[] box) -> Tensor",
    "pass1": "xint::pass1(Tensor sb, Tensor?[] rows, Tensor?[] det, float[] det_scalars, float[] call, float[]? rcw, "
             "Tensor? acc_num=None, Tensor? acc_den=None) -> Tensor[]",
    "accumulate_profiles": "xint::accumulate_profiles(Tensor img2, Tensor ok2, Tensor J, Tensor u1, Tensor u2, "
                           "Tensor region, Tensor isum, Tensor sel, Tensor seff, int n_regions) -> (Tensor, Tensor)",
    "normalise_profiles": "xint::normalise_profiles(Tensor num_fx, Tensor den_fx, Tensor gauss) "
                          "-> (Tensor, Tensor, Tensor)",
    "pass2": "xint::pass2(Tensor sb, Tensor?[] rows, Tensor?[] det, float[] det_scalars, float[] call, float[]? rcw, "
             "Tensor prof, Tensor bg, Tensor nbg) -> Tensor[]",
}


def test_operators_registered_with_their_schemas():
    from xpg_amd import ops
    xint = ops.load()
    for name, schema in SCHEMAS.items():
        assert str(getattr(xint, name).default._schema) == schema, name


def test_host_tensors_are_refused():
    from xpg_amd import ops
    xint = ops.load()
    sb = torch.zeros((2, 3, 5, 5), dtype=torch.float32)
    with pytest.raises((RuntimeError, NotImplementedError)):
        xint.pass1(sb, [None] * 16, [None] * 5, [0.0] * 8, [0.0] * 11, None)

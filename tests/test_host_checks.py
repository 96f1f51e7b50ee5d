"""Host-side helpers that GPU tests and the benchmark use as float64 checks,
pinned on the CPU: the sampled convolution against the full float64
convolution (which is itself a closed form over numpy), and bf16 rounding
against torch's float32 -> bfloat16 conversion."""
import numpy as np

from loop_tool_amd import programs as P


def test_sampled_conv_matches_full_conv():
    X = P.normal(4, (3, 5, 9, 7))
    W = P.normal(5, (4, 5, 3, 3))
    full = P.reference_output("conv2d", [X, W])
    idx = P.conv_sample_indices(3, 4, 9, 7, n_random=500)
    # every border pixel of planes n=0 and n=2, for every output channel
    border = (idx[:, 2] == 0) | (idx[:, 2] == 8) | (idx[:, 3] == 0) | (idx[:, 3] == 6)
    assert {tuple(r) for r in idx[border][:, :2]} >= {(n, c) for n in (0, 2) for c in range(4)}
    assert border[:2 * 4 * (2 * 9 + 2 * 7 - 4)].all()
    got = P.conv_f64_at(X, W, idx, chunk=64)
    want = full[idx[:, 0], idx[:, 1], idx[:, 2], idx[:, 3]]
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12)
    # a window shifted by one pixel (no padding) is caught
    shifted = P.conv_f64_at(np.pad(X, ((0, 0), (0, 0), (0, 2), (0, 2))), W, idx, pad=0,
                            chunk=64)
    assert not np.allclose(shifted, want)


def test_sampled_conv_catches_a_transposed_weight():
    X = P.normal(4, (2, 6, 8, 8))
    W = P.normal(5, (6, 6, 3, 3))
    idx = P.conv_sample_indices(2, 6, 8, 8, n_random=200)
    full = P.reference_output("conv2d", [X, W.transpose(1, 0, 2, 3)])
    got = P.conv_f64_at(X, W, idx)
    assert not np.allclose(got, full[idx[:, 0], idx[:, 1], idx[:, 2], idx[:, 3]])


def test_bf16_rounding_matches_torch():
    import torch
    x = np.concatenate([P.normal(3, (1 << 16,), 3.0),
                        np.array([1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -0.0, 65504.0], np.float32)])
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(P.round_bf16(x).view(np.uint32), want.view(np.uint32))

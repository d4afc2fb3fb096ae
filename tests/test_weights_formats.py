"""Model-file formats of the reference (python/convert.py:41-57,
src/HignnModel.cpp:404-411): torch.save of the Sequential (.pkl) and
torch.jit.save of Net (.pt, base_<device>.pt), read without unpickling."""
import os
import pickle
import zipfile

import numpy as np
import pytest

from hignn_amd import weights

torch = pytest.importorskip("torch")


def _sequential(w):
    parts = weights.unpack(w)
    seq = torch.nn.Sequential(torch.nn.Linear(3, 128), torch.nn.Tanh(), torch.nn.Linear(128, 512),
                              torch.nn.Tanh(), torch.nn.Linear(512, 9))
    with torch.no_grad():
        for i, (wn, bn) in zip([0, 2, 4], [("W1", "b1"), ("W2", "b2"), ("W3", "b3")]):
            seq[i].weight.copy_(torch.from_numpy(parts[wn]))
            seq[i].bias.copy_(torch.from_numpy(parts[bn]))
    return seq


class _Net(torch.nn.Module):
    """The wrapper of python/convert.py:6-20 (+I, /|r|)."""

    def __init__(self, model):
        super().__init__()
        self.model = model

    def forward(self, x):
        r = torch.norm(x, dim=1, keepdim=True)
        y = self.model(x)
        y[:, 0] += 1
        y[:, 4] += 1
        y[:, 8] += 1
        return y / r


def test_torch_save_pkl(tmp_path):
    w = weights.xavier(3)
    torch.save(_sequential(w), tmp_path / "m.pkl")
    np.testing.assert_array_equal(weights.load(str(tmp_path / "m")), w)


def test_torchscript_pt_and_device_suffix(tmp_path):
    w = weights.xavier(4)
    sm = torch.jit.script(_Net(_sequential(w)))
    torch.jit.save(sm, str(tmp_path / "m_0.pt"))
    torch.jit.save(sm, str(tmp_path / "n.pt"))
    assert weights.resolve_model_path(str(tmp_path / "m"), device=0).endswith("m_0.pt")
    np.testing.assert_array_equal(weights.load(str(tmp_path / "m")), w)
    np.testing.assert_array_equal(weights.load(str(tmp_path / "n")), w)


def test_loader_invokes_nothing(tmp_path):
    """A data.pkl whose objects would run code when unpickled: the reader
    only walks the opcodes, so nothing runs, and it reports the missing
    tensors."""
    marker = tmp_path / "ran"

    class Evil:
        def __reduce__(self):
            return (exec, (f"open({str(marker)!r}, 'w').write('x')",))

    arc = tmp_path / "evil.pkl"
    with zipfile.ZipFile(arc, "w") as z:
        z.writestr("evil/data.pkl", pickle.dumps({"model": Evil()}, protocol=2))
    with pytest.raises(ValueError, match="0 of the 6 MLP tensors"):
        weights.load_torch_zip(str(arc))
    assert not marker.exists()


def test_missing_model(tmp_path):
    with pytest.raises(FileNotFoundError):
        weights.load(str(tmp_path / "nothing"))

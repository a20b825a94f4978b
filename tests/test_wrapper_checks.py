"""CPU checks of the Python binding's argument checks: every tensor is
validated (GPU, dtype, contiguous, shape against the context) before the
library is called, so a wrong argument raises ValueError on the host rather
than reaching a kernel as an out-of-bounds address."""
import pytest
import torch

from code_amd import transport as gt


class _NoLibrary:
    def __getattr__(self, name):
        def call(*a):
            raise AssertionError(f"the library was called ({name}) despite a bad argument")
        return call


def fake_context(imax=10, jmax=8, kmax=4):
    c = object.__new__(gt.Context)
    c.lib = _NoLibrary()
    c.h = None
    c.imin = c.jmin = 1
    c.imax, c.jmax, c.kmax = imax, jmax, kmax
    c.NX, c.NY = imax + 4, jmax + 4
    return c


def test_cpu_tensor_rejected():
    c = fake_context()
    f = torch.zeros(c.shape3, dtype=torch.float64)
    with pytest.raises(ValueError, match="GPU"):
        c.do_advection_3d(30.0, f, f, f, f, f, f, f, f, gt.HALFSPLIT, gt.P2_PDM, gt.P2_PDM,
                          -1.0, gt.H_TAG)


def test_dev_checks_dtype_shape_contiguity(monkeypatch):
    # pretend the tensors live on the GPU: only the other checks remain
    monkeypatch.setattr(torch.Tensor, "device", property(lambda self: torch.device("cuda")))
    good = torch.zeros((5, 12, 14), dtype=torch.float64)
    assert gt._dev(good, "f", shape=(5, 12, 14)).value == good.data_ptr()
    with pytest.raises(ValueError, match="dtype"):
        gt._dev(good.float(), "f", shape=(5, 12, 14))
    with pytest.raises(ValueError, match="shape"):
        gt._dev(good[:, :, :13].contiguous(), "f", shape=(5, 12, 14))
    with pytest.raises(ValueError, match="contiguous"):
        gt._dev(good.transpose(1, 2), "f")
    with pytest.raises(ValueError, match="elements"):
        gt._dev(good, "f", numel=7)
    with pytest.raises(ValueError, match="required"):
        gt._dev(None, "f")
    assert gt._dev(None, "f", optional=True) is None
    with pytest.raises(ValueError, match="torch tensor"):
        gt._dev(good.numpy(), "f")


def test_context_methods_check_against_the_domain(monkeypatch):
    monkeypatch.setattr(torch.Tensor, "device", property(lambda self: torch.device("cuda")))
    c = fake_context()
    f3 = torch.zeros(c.shape3, dtype=torch.float64)
    f2 = torch.zeros(c.shape2, dtype=torch.float64)
    wrong = torch.zeros((c.kmax + 2,) + c.shape2, dtype=torch.float64)
    with pytest.raises(ValueError, match="shape"):
        c.do_advection_3d(30.0, wrong, f3, f3, f3, f3, f3, f3, f3, 2, 6, 6, -1.0, 10)
    with pytest.raises(ValueError, match="shape"):
        c.adv_split_u(30.0, f3, f3, f3, f3, f3, f3, 0.5, 6, -1.0)   # slices are 2-D
    with pytest.raises(ValueError, match="shape"):
        c.ip_blumberg_mellor(f3, f3, f3, f3, f3, f3, f3, f3)      # H is 2-D
    with pytest.raises(ValueError, match="dtype"):
        c.budget(f3, f3.to(torch.float16))
    with pytest.raises(ValueError, match="required"):
        c.do_internal_pressure(6, f2, f3, f3, f3, f3, None, f3, f3)  # SMcW reads sseo
    with pytest.raises(ValueError, match="elements"):
        c.adv3d_adjoint_f32(3, 4, 2, 30.0, f3.float(), f3.float(), f3.float(), f3, f3, f3,
                            f3, f3, f3, 4, 4)


def test_batch_checks(monkeypatch):
    monkeypatch.setattr(torch.Tensor, "device", property(lambda self: torch.device("cuda")))
    c1, c2 = fake_context(), fake_context(12, 6, 4)

    class H:
        value = 1
    c1.h = c2.h = H()

    def dom(c, ntr, dtype=torch.float32):
        f3 = torch.zeros(c.shape3, dtype=torch.float64)
        return dict(ctx=c, f_in=torch.zeros((ntr,) + c.shape3, dtype=dtype),
                    f_out=torch.ones((ntr,) + c.shape3, dtype=dtype), uu=f3, vv=f3, ww=f3,
                    hu=f3, hv=f3, ho=f3)
    assert len(gt._batch_array([dom(c1, 3), dom(c2, 3)], 3)) == 2
    with pytest.raises(ValueError, match="repeats a context"):
        gt._batch_array([dom(c1, 3), dom(c1, 3)], 3)
    with pytest.raises(ValueError, match="dtype"):
        gt._batch_array([dom(c1, 3), dom(c2, 3, torch.float64)], 3)
    with pytest.raises(ValueError, match="elements"):
        gt._batch_array([dom(c1, 2)], 3)          # fewer tracers than ntracers
    d = dom(c1, 3)
    d["f_out"] = d["f_in"]
    with pytest.raises(ValueError, match="differ"):
        gt._batch_array([d], 3)

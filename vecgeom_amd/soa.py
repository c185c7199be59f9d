"""SOA3D: three contiguous float64 device arrays x, y, z (VecGeom/base/SOA3D.h:35-41), and the
check every wrapper applies to a tensor before its address is handed to the C ABI."""
import torch


def device_array(t, dtypes, n, name):
    """Returns t if it is a contiguous device tensor of one of `dtypes` with exactly n elements;
    raises ValueError otherwise (the kernels read and write n elements through its address)."""
    dtypes = dtypes if isinstance(dtypes, tuple) else (dtypes,)
    if not torch.is_tensor(t):
        raise ValueError("%s must be a torch tensor" % name)
    if not t.is_cuda:
        raise ValueError("%s must be a device tensor" % name)
    if t.dtype not in dtypes:
        raise ValueError("%s must be %s, not %s" % (name, " or ".join(str(d) for d in dtypes), t.dtype))
    if not t.is_contiguous():
        raise ValueError("%s must be contiguous" % name)
    if t.numel() != n:
        raise ValueError("%s has %d elements, %d expected" % (name, t.numel(), n))
    return t


def same_length(points, dirs):
    if dirs is not None and len(dirs) != len(points):
        raise ValueError("directions (%d) and points (%d) differ in length" % (len(dirs), len(points)))


class SOA3D:
    def __init__(self, x, y, z):
        for a in (x, y, z):
            if a.dtype != torch.float64 or not a.is_cuda or not a.is_contiguous():
                raise ValueError("SOA3D components must be contiguous float64 device tensors")
        if not (x.numel() == y.numel() == z.numel()):
            raise ValueError("SOA3D components differ in length")
        self.x, self.y, self.z = x, y, z

    @classmethod
    def from_array(cls, a, device="cuda"):
        """a: array-like of shape (3, n)."""
        t = torch.as_tensor(a, dtype=torch.float64)
        return cls(*(t[i].contiguous().to(device) for i in range(3)))

    @classmethod
    def empty(cls, n, device="cuda"):
        return cls(*(torch.empty(n, dtype=torch.float64, device=device) for _ in range(3)))

    @classmethod
    def zeros(cls, n, device="cuda"):
        return cls(*(torch.zeros(n, dtype=torch.float64, device=device) for _ in range(3)))

    def size(self):
        return self.x.numel()

    def __len__(self):
        return self.x.numel()

    def ptrs(self):
        return self.x.data_ptr(), self.y.data_ptr(), self.z.data_ptr()

    def cpu(self):
        return torch.stack([self.x, self.y, self.z]).cpu().numpy()

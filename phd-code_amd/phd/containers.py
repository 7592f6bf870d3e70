"""Named structure-of-arrays containers backed by torch tensors.

CarrayContainer mirrors the reference's container API
(containers/containers.pyx:13-395): named typed arrays of one common length,
named groups of field names, resize / extract / append / remove-by-tag.
Arrays live on the GPU (cuda:0) when one is visible, else on the host; the
reference's dtype names keep their reference meaning (utils/carray.pxd):
"double" float64, "int" int8, "long" int32, "longlong" int64.

GradientContainer holds the per-cell primitive gradients cell-major (one row
per cell, row length padded to an even count so rows are 16-byte aligned), and
exposes each component "<field>_<axis>" as a strided column view, with the
reference's named groups (reconstruction.pyx:462-509).
"""
import torch

DTYPES = {"double": torch.float64, "int": torch.int8, "long": torch.int32,
          "longlong": torch.int64}


def default_device():
    return torch.device("cuda", torch.cuda.current_device()) if torch.cuda.is_available() \
        else torch.device("cpu")


class CarrayContainer(object):
    def __init__(self, num=0, carrays_to_register=None, device=None):
        self.device = default_device() if device is None else torch.device(device)
        self.carrays = {}
        self.carray_dtypes = {}
        self.carray_named_groups = {}
        self._size = int(num)
        self._cap = int(num)
        if carrays_to_register:
            for name, dtype in carrays_to_register.items():
                self.register_carray(num, name, dtype)

    # -- registration -------------------------------------------------------
    def register_carray(self, carray_size, name, dtype="double"):
        if name in self.carrays:
            raise ValueError("Carray %s already registered" % name)
        if self.carrays and carray_size != self._size:
            raise ValueError("Size inconsistent with carray size")
        if dtype not in DTYPES:
            raise ValueError("unknown carray dtype %r" % dtype)
        if not self.carrays:
            self._size = self._cap = int(carray_size)
        self.carrays[name] = torch.zeros(self._cap, dtype=DTYPES[dtype], device=self.device)
        self.carray_dtypes[name] = dtype

    @property
    def carrays_dtypes(self):
        return self.carray_dtypes

    # -- access -------------------------------------------------------------
    def get_carray_size(self):
        return self._size

    def get_carray(self, name):
        return self.carrays[name][:self._size]

    def __getitem__(self, name):
        return self.carrays[name][:self._size]

    def __setitem__(self, name, value):
        self.carrays[name][:self._size] = value

    def __contains__(self, name):
        return name in self.carrays

    def properties(self):
        return dict(self.carray_dtypes)

    def group(self, name):
        """Tensors of a named group, in group order."""
        return [self[f] for f in self.carray_named_groups[name]]

    # -- size changes -------------------------------------------------------
    def resize(self, size):
        size = int(size)
        if size > self._cap:
            cap = max(size, int(self._cap * 1.25) + 16)
            for name, t in self.carrays.items():
                new = torch.zeros(cap, dtype=t.dtype, device=self.device)
                new[:self._size] = t[:self._size]
                self.carrays[name] = new
            self._cap = cap
        self._size = size

    def extend(self, num):
        self.resize(self._size + int(num))

    def remove_items(self, indices):
        keep = torch.ones(self._size, dtype=torch.bool, device=self.device)
        keep[torch.as_tensor(indices, device=self.device, dtype=torch.long)] = False
        self._compact(keep)

    def remove_tagged_particles(self, tag):
        self._compact(self["tag"] != tag)

    def _compact(self, keep):
        n = int(keep.sum().item())
        for name, t in self.carrays.items():
            t[:n] = t[:self._size][keep]
        self._size = n

    def extract_items(self, indices, fields=None):
        idx = torch.as_tensor(indices, device=self.device, dtype=torch.long)
        names = list(self.carrays) if fields is None else list(fields)
        out = CarrayContainer(device=self.device)
        out._size = out._cap = int(idx.numel())
        for name in names:
            out.carrays[name] = self.carrays[name][:self._size][idx].clone()
            out.carray_dtypes[name] = self.carray_dtypes[name]
        out.carray_named_groups = {k: list(v) for k, v in self.carray_named_groups.items()}
        return out

    def append_container(self, other):
        n0 = self._size
        self.resize(n0 + other.get_carray_size())
        for name in self.carrays:
            if name in other.carrays:
                self.carrays[name][n0:self._size] = other[name]

    def copy(self):
        return self.extract_items(torch.arange(self._size, device=self.device))

    def pointer(self, name):
        return self.carrays[name].data_ptr()

    def pointer_groups(self, group):
        return [self.carrays[f].data_ptr() for f in self.carray_named_groups[group]]


class GradientContainer(object):
    """Cell-record rows holding the gradients.

    Row layout (one row per cell): primitive(D+2) | gradient(D(D+2)) |
    centroid(D), padded to an even length (24 doubles in 3D, 14 in 2D), so
    the face kernel reads everything it needs about a cell in one
    contiguous row.  Gradient component "<field>_<axis>" (column D+2 +
    n*D + k, d prim_n / d x_k) is exposed as a strided column view with the
    reference's named groups (reconstruction.pyx:462-509)."""

    def __init__(self, field_names, dim, ld, device=None):
        self.device = default_device() if device is None else torch.device(device)
        self.dim = dim
        axis = "xyz"
        self.fields = list(field_names)
        self.names = []
        groups = {"velocity": [], "primitive": []}
        for f in self.fields:
            groups[f] = []
        for f in self.fields:
            for k in range(dim):
                g = f + "_" + axis[k]
                self.names.append(g)
                groups["primitive"].append(g)
                groups[f].append(g)
                if "vel" in f:
                    groups["velocity"].append(g)
        self.carray_named_groups = groups
        self.offset = len(self.fields)               # gradient columns follow the primitives
        self.column = {g: self.offset + c for c, g in enumerate(self.names)}
        self.ld = int(ld)
        self.data = torch.zeros((0, self.ld), dtype=torch.float64, device=self.device)
        self._size = 0

    @property
    def carrays(self):
        return {g: self[g] for g in self.names}

    def get_carray_size(self):
        return self._size

    def resize(self, n):
        n = int(n)
        if n > self.data.shape[0]:
            new = torch.zeros((max(n, int(self.data.shape[0] * 1.25) + 16), self.ld),
                              dtype=torch.float64, device=self.device)
            new[:self._size] = self.data[:self._size]
            self.data = new
        self._size = n

    def __getitem__(self, name):
        return self.data[:self._size, self.column[name]]

    def __setitem__(self, name, value):
        self.data[:self._size, self.column[name]] = value

    def rows(self):
        """Gradient components of every cell, [n_cells, D(D+2)]."""
        return self.data[:self._size, self.offset:self.offset + len(self.names)]

"""Snapshot writers with the reference's Hdf5 tree (io/read_write.py:1-100).

The tree is the reference's:

    attrs: dt, time, iteration
    particles/                 attrs: Real, Ghost, number_particles
    particles/<field>/data     one array per carray, in registration order
    particles/<field>          attrs: dtype ("double", "int", "long", "longlong")

Hdf5 writes it with h5py when h5py is importable.  NpzTree writes the same
tree into one .npz archive (no pickling: attributes are stored as arrays
under "<group>/attrs/<name>") so snapshots work where h5py is absent.  Field
arrays are copied from the GPU once per write.
"""
import os

import numpy as np
import torch

from .containers import CarrayContainer
from .utils import ParticleTAGS


class ReaderWriterBase(object):
    def write(self, base_name, output_directory, integrator):
        raise NotImplementedError("ReaderWriterBase::write called!")

    def read(self, file_name):
        raise NotImplementedError("ReaderWriterBase::read called!")


def _host(t):
    return t.detach().cpu().numpy()


def _tree(integrator):
    """The reference's snapshot tree as flat path -> numpy value pairs."""
    pc = integrator.particles
    out = {
        "attrs/dt": np.float64(integrator.dt),
        "attrs/time": np.float64(integrator.time),
        "attrs/iteration": np.int64(integrator.iteration),
        "particles/attrs/Real": np.int64(ParticleTAGS.Real),
        "particles/attrs/Ghost": np.int64(ParticleTAGS.Ghost),
        "particles/attrs/number_particles": np.int64(pc.get_carray_size()),
    }
    for name in pc.carrays.keys():
        out["particles/%s/attrs/dtype" % name] = np.array(pc.carray_dtypes[name])
        out["particles/%s/data" % name] = _host(pc[name])
    return out


def _container(num, fields, device=None):
    """fields: list of (name, dtype, data) in file order."""
    particles = CarrayContainer(int(num), device=device)
    for name, dtype, data in fields:
        particles.register_carray(int(num), name, dtype)
        particles[name][:] = torch.as_tensor(np.asarray(data), device=particles.device)
    return particles


class NpzTree(ReaderWriterBase):
    """The Hdf5 tree in a .npz archive."""

    def write(self, base_name, output_directory, integrator):
        path = os.path.join(output_directory, base_name + ".npz")
        np.savez(path, **_tree(integrator))
        return path

    def read(self, file_name, device=None):
        with np.load(file_name, allow_pickle=False) as z:
            num = int(z["particles/attrs/number_particles"])
            names = [k[len("particles/"):-len("/data")] for k in z.files
                     if k.startswith("particles/") and k.endswith("/data")]
            fields = [(n, str(z["particles/%s/attrs/dtype" % n]), z["particles/%s/data" % n])
                      for n in names]
        return _container(num, fields, device)


class Hdf5(ReaderWriterBase):
    """io/read_write.py:53-100, written with h5py."""

    def _h5py(self):
        try:
            import h5py
        except ImportError:
            raise RuntimeError("Hdf5 snapshots need h5py, which is not installed; "
                               "NpzTree writes the same tree without it")
        return h5py

    def write(self, base_name, output_directory, integrator):
        h5py = self._h5py()
        path = os.path.join(output_directory, base_name + ".hdf5")
        tree = _tree(integrator)
        with h5py.File(path, "w") as f:
            for k in ("dt", "time", "iteration"):
                f.attrs[k] = tree["attrs/" + k]
            g = f.create_group("particles")
            for k in ("Real", "Ghost", "number_particles"):
                g.attrs[k] = tree["particles/attrs/" + k]
            for name in integrator.particles.carrays.keys():
                d = g.create_group(name)
                d.attrs["dtype"] = str(tree["particles/%s/attrs/dtype" % name])
                d.create_dataset("data", data=tree["particles/%s/data" % name])
        return path

    def read(self, file_name, device=None):
        h5py = self._h5py()
        with h5py.File(file_name, "r") as f:
            g = f["particles"]
            num = int(g.attrs["number_particles"])
            fields = [(n, str(g[n].attrs["dtype"]), g[n]["data"][:]) for n in g.keys()]
        return _container(num, fields, device)

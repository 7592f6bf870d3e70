"""Builders of the C-ABI views (phd_cells / phd_faces) over containers."""
from . import _native as nat


def cells_view(particles):
    """phd_cells over a particle container (reals first, then ghosts)."""
    g = particles.carray_named_groups
    dim = len(g["position"])
    c = nat.Cells()
    c.dim = dim
    c.n_cells = particles.get_carray_size()
    c.tag = particles.pointer("tag")
    c.type = particles.pointer("type")
    c.map = particles.pointer("map") if "map" in particles.carrays else None
    for k, p in enumerate(particles.pointer_groups("position")):
        c.x[k] = p
    if "dcom" in g:
        for k, p in enumerate(particles.pointer_groups("dcom")):
            c.dcom[k] = p
    if "volume" in particles.carrays:
        c.volume = particles.pointer("volume")
    for q, p in enumerate(particles.pointer_groups("primitive")):
        c.prim[q] = p
    for q, p in enumerate(particles.pointer_groups("conservative")):
        c.cons[q] = p
    if "w" in g:
        for k, p in enumerate(particles.pointer_groups("w")):
            c.w[k] = p
    return c


def faces_view(mesh):
    """phd_faces over a mesh's face container and per-cell face lists."""
    f = nat.Faces()
    faces = mesh.faces
    f.n_faces = faces.get_carray_size()
    f.area = faces.pointer("area")
    for k, p in enumerate(faces.pointer_groups("com")):
        f.com[k] = p
    for k, p in enumerate(faces.pointer_groups("normal")):
        f.normal[k] = p
    for k, p in enumerate(faces.pointer_groups("velocity")):
        f.velocity[k] = p
    f.pair_i = faces.pointer("pair-i")
    f.pair_j = faces.pointer("pair-j")
    f.nbr_off = mesh.neighbors.offsets.data_ptr()
    f.nbr_face = mesh.neighbors.faces.data_ptr() if mesh.neighbors.faces.numel() else 0
    return f

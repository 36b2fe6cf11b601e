"""Material-index grid: a uint16 id per cell plus a small table of distinct materials.

The reference keeps float32 material arrays per cell (core/material_grid.py:38-83) and rebuilds per-cell coefficient
arrays from them (physics/em/solvers/fdtd_ez_2d_kernels.py:310-353); its 3D engine already stores a uint16 id per cell
with per-id tables (physics/em3d/material_index3d.py:131-179). Here every grid uses the id form: boxes are painted as
index ranges (later writes win, as material_grid.py:56-66), material values are float32 as in the reference, and the
update coefficients of each id are computed in float64 and rounded once to float32.
"""
import numpy as np

from .constants import DTYPE, EPS0, MU0

MAX_MATERIALS = 65535
# The perfect electric conductor: a material entry with eps = -1 on every axis, mu_r = 1 and no losses (the reference's
# last table id, physics/em3d/engine3d.py:339-362). Its E coefficients are pinned to 0 in every table; a face between
# a PEC cell and another cell follows the rule of each polarisation (face coefficients in the kernels).
PEC_VALUES = (-1.0, -1.0, -1.0, 0.0, 1.0, 0.0)


def semi_implicit(eps_rel, sigma, dt, vacuum):
    """(ca, cb) of eps dE/dt = curl H - sigma E with eps = eps_rel * vacuum: k = sigma dt / (2 eps),
    ca = (1 - k)/(1 + k), cb = (dt / eps)/(1 + k) (fdtd_ez_2d_kernels.py:310-353), float64."""
    eps = np.asarray(eps_rel, dtype=np.float64) * vacuum
    k = np.asarray(sigma, dtype=np.float64) * dt / (2.0 * eps)
    denom = 1.0 + k
    return (1.0 - k) / denom, (dt / eps) / denom


class MaterialMap:
    """uint16 id per cell; materials[id] = (eps_x, eps_y, eps_z, sigma_e, mu_r, sigma_m) as float32 values."""

    def __init__(self, shape):
        self.shape = tuple(shape)
        self.ids = np.zeros(self.shape, dtype=np.uint16)
        self.materials = [(DTYPE(1.0),) * 3 + (DTYPE(0.0), DTYPE(1.0), DTYPE(0.0))]
        self._lookup = {self.materials[0]: 0}
        self.version = 0
        # the dielectric ids painted under PEC cells (None until PEC is painted): the reference keeps PEC in its
        # engine's own index and its material arrays hold the dielectric (physics/em3d/engine3d.py:351-362)
        self.dielectric_ids = None

    def material_id(self, eps_x, eps_y=None, eps_z=None, sigma_e=0.0, mu_r=1.0, sigma_m=0.0):
        eps_y = eps_x if eps_y is None else eps_y
        eps_z = eps_x if eps_z is None else eps_z
        key = tuple(DTYPE(v) for v in (eps_x, eps_y, eps_z, sigma_e, mu_r, sigma_m))
        if not all(np.isfinite(key)):
            raise ValueError(f'material values must be finite, got {key}')
        if key[0] <= 0 or key[1] <= 0 or key[2] <= 0 or key[4] <= 0:
            raise ValueError(f'permittivity and permeability must be > 0, got {key}')
        index = self._lookup.get(key)
        if index is None:
            if len(self.materials) >= MAX_MATERIALS:
                raise ValueError(f'more than {MAX_MATERIALS} distinct materials')
            index = len(self.materials)
            self.materials.append(key)
            self._lookup[key] = index
        return index

    def pec_id(self):
        """Id of the PEC entry (created on first use)."""
        key = tuple(DTYPE(v) for v in PEC_VALUES)
        index = self._lookup.get(key)
        if index is None:
            if len(self.materials) >= MAX_MATERIALS:
                raise ValueError(f'more than {MAX_MATERIALS} distinct materials')
            index = len(self.materials)
            self.materials.append(key)
            self._lookup[key] = index
        return index

    def pec_flags(self):
        """bool (n_materials,): the PEC entries."""
        return self.values()[:, 0] < 0.0

    def paint_box(self, ranges, material):
        """Write `material` into the half-open index box `ranges` (one (lo, hi) per axis)."""
        if any(hi <= lo for lo, hi in ranges):
            return
        self.ids[tuple(slice(lo, hi) for lo, hi in ranges)] = material
        self.version += 1

    def paint_mask(self, mask, material):
        self.ids[mask] = material
        self.version += 1

    def paint_pec(self, keys):
        """Paint PEC over the cells of every index expression in `keys` (slice tuples or bool masks), keeping the
        dielectric painted under them in dielectric_ids; returns whether a cell changed."""
        pec = self.pec_id()
        if not any(np.any(self.ids[key] != pec) for key in keys):
            return False
        under = self.ids.copy() if self.dielectric_ids is None else self.dielectric_ids
        fresh = self.ids != pec   # cells painted since the last PEC pass hold the newest dielectric
        under[fresh] = self.ids[fresh]
        self.dielectric_ids = under
        for key in keys:
            self.ids[key] = pec
        self.version += 1
        return True

    def display_ids(self, key=Ellipsis):
        """The ids of the cells `key` with every PEC cell showing the dielectric painted under it: the material grid
        the reference's renderers show (its PEC lives only in the engine's index)."""
        ids = self.ids[key]
        if self.dielectric_ids is None:
            return ids
        return np.where(self.pec_flags()[ids], self.dielectric_ids[key], ids)

    def display_array(self, column, key=Ellipsis):
        """Per-cell float32 values of one material column as displayed (display_ids)."""
        table = np.array([m[column] for m in self.materials], dtype=DTYPE)
        return table[self.display_ids(key)]

    def set_property(self, columns, values):
        """Write the per-cell float32 array `values` into material columns `columns` (e.g. (0, 1, 2) for an isotropic
        eps_r, (3,) for sigma_e) of every cell, as a direct write into the reference's material arrays followed by
        materials_changed does (core/material_grid.py:38-83). Each distinct (old material, new value) pair becomes one
        id; PEC cells stay PEC, and the dielectric under them takes the values too."""
        values = np.broadcast_to(np.asarray(values, dtype=DTYPE), self.shape)
        new_ids = self._with_property(self.ids, columns, values)
        if self.dielectric_ids is not None:
            self.dielectric_ids = self._with_property(self.dielectric_ids, columns, values)
        self.ids = new_ids
        self.version += 1

    def _with_property(self, ids, columns, values):
        """ids remapped so that each cell's material takes `values` in `columns` (PEC entries kept)."""
        key = (ids.astype(np.uint64) << np.uint64(32)) | values.view(np.uint32).astype(np.uint64)
        unique, inverse = np.unique(key.ravel(), return_inverse=True)
        pec = self.pec_flags()
        mapped = np.empty(unique.size, dtype=np.uint16)
        for u, packed in enumerate(unique):
            old = int(packed >> np.uint64(32))
            if pec[old]:
                mapped[u] = old
                continue
            value = np.array([int(packed & np.uint64(0xFFFFFFFF))], dtype=np.uint32).view(DTYPE)[0]
            entry = list(self.materials[old])
            for column in columns:
                entry[column] = value
            mapped[u] = self.material_id(*entry)
        return mapped[inverse].reshape(self.shape)

    def values(self):
        """(n_materials, 6) float64 array of the float32 material values."""
        return np.array(self.materials, dtype=np.float64)

    def used(self):
        """Sorted unique ids present in the grid."""
        return np.unique(self.ids)

    def property_array(self, column):
        """Per-cell float32 array of one material column (0 eps_x, 1 eps_y, 2 eps_z, 3 sigma_e, 4 mu_r, 5 sigma_m)."""
        table = np.array([m[column] for m in self.materials], dtype=DTYPE)
        return table[self.ids]

    # kernel tables ---------------------------------------------------------------------------------------------
    def table_ez(self, dt):
        """(n, 4) float32 (ca, cb, da, db) per id (E_z polarisation)."""
        v = self.values()
        dt = float(dt)
        ca, cb = semi_implicit(v[:, 0], v[:, 3], dt, EPS0)
        da, db = semi_implicit(v[:, 4], v[:, 5], dt, MU0)
        pec = self.pec_flags()
        ca[pec] = cb[pec] = 0.0   # E_z pinned to 0 in PEC cells (physics/em/solvers/fdtd_ez_2d.py:223-233)
        return np.stack([ca, cb, da, db], axis=1).astype(DTYPE)

    def tables_hz(self, dt):
        """(n, 4) float32 (da, db, eps_xx, eps_yy) and (n, 8) float32 (ca_x, cb_x, ca_y, cb_y, sigma_e, 0, 0, 0):
        the face coefficients between two cells of the same id (H_z polarisation)."""
        v = self.values()
        dt = float(dt)
        da, db = semi_implicit(v[:, 4], v[:, 5], dt, MU0)
        ca_x, cb_x = semi_implicit(v[:, 0], v[:, 3], dt, EPS0)
        ca_y, cb_y = semi_implicit(v[:, 1], v[:, 3], dt, EPS0)
        pec = self.pec_flags()
        for column in (ca_x, cb_x, ca_y, cb_y):
            column[pec] = 0.0   # every E face touching a PEC cell is pinned (fdtd_hz_2d.py:475-495)
        table = np.stack([da, db, v[:, 0], v[:, 1]], axis=1).astype(DTYPE)
        zeros = np.zeros_like(da)
        table2 = np.stack([ca_x, cb_x, ca_y, cb_y, v[:, 3], zeros, zeros, zeros], axis=1).astype(DTYPE)
        return table, table2

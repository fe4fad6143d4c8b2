"""Fully threaded oct tree in the reference's memory model.

The arrays are those of amr/amr_commons.f90:54-75 with the reference's
1-based index conventions, stored in 0-based numpy arrays:

* cell index of cell ``ind`` (1..8) of grid ``igrid`` (1..ngridmax) is
  ``ncoarse + (ind-1)*ngridmax + igrid`` (hydro/godunov_fine.f90:601), with
  ``ind = 1 + ix + 2*iy + 4*iz`` for the cell's position in its oct;
* ``son[icell-1]``    grid index of the cell's child oct, 0 for a leaf;
* ``father[igrid-1]`` index of the oct's father cell;
* ``nbor[(j-1)*ngridmax + igrid-1]``, j = 1..6 (-x,+x,-y,+y,-z,+z): the
  neighbour cells of the father cell (nbor(igrid,j));
* ``active[l]``       the grids of level l in sweep order.

The coarse grid is one cell (nx=ny=nz=1, boxlen=1, periodic), so level-l
cells have size 2^-l.  The tree is built from nested refinement regions: a
level-l cell is refined when the caller's criterion says so, dilated by
``nexpand`` cells, and every refined cell's 3x3x3 neighbourhood is made to
exist by refining the coarser level around it -- the rule that lets
get3cubefather (amr/nbors_utils.f90:5-194) find all 27 father neighbours.
"""
import numpy as np

NCOARSE = 1


def _dilate(mask, r):
    """Periodic dilation of a boolean cube by r cells in the 3x3x3 sense."""
    out = mask.copy()
    for _ in range(r):
        m = out.copy()
        for ax in range(3):
            m = m | np.roll(m, 1, axis=ax) | np.roll(m, -1, axis=ax)
        out = m
    return out


class OctTree:
    def __init__(self, levelmin, levelmax, refine=None, nexpand=1, extra_grids=0):
        """refine(level, centers) -> bool array for the level's cells
        (centers: (N, 3) positions in [0,1)); consulted for levelmin <= level
        < levelmax.  Levels below levelmin are fully refined."""
        if not 1 <= levelmin <= levelmax <= 8:
            raise ValueError("this builder supports 1 <= levelmin <= levelmax <= 8")
        self.levelmin, self.levelmax = levelmin, levelmax
        # refined-cell masks per level on the full lattice (index [z][y][x])
        R = {}
        for l in range(1, levelmax):
            n = 2 ** l
            if l < levelmin:
                R[l] = np.ones((n, n, n), dtype=bool)
            elif refine is None:
                R[l] = np.zeros((n, n, n), dtype=bool)
            else:
                c = (np.indices((n, n, n)).reshape(3, -1)[::-1].T + 0.5) / n  # (x,y,z)
                R[l] = np.asarray(refine(l, c), dtype=bool).reshape(n, n, n)
        for l in range(levelmax - 1, levelmin - 1, -1):
            R[l] = _dilate(R[l], nexpand)
            if l - 1 >= 1:
                need = _dilate(R[l], 1)  # 3^3 neighbourhood must exist at level l
                n = 2 ** (l - 1)
                parent = need.reshape(n, 2, n, 2, n, 2).any(axis=(1, 3, 5))
                R[l - 1] = R[l - 1] | parent
        self.R = R
        # grids level by level, children in (parent order, ind order)
        levels = {1: [(0, 0, 0)]}
        for l in range(1, levelmax):
            kids = []
            for (gx, gy, gz) in levels[l]:
                for ind in range(8):
                    ix, iy, iz = ind & 1, (ind >> 1) & 1, ind >> 2
                    cx, cy, cz = 2 * gx + ix, 2 * gy + iy, 2 * gz + iz
                    if R[l][cz, cy, cx]:
                        kids.append((cx, cy, cz))
            if not kids:
                break
            levels[l + 1] = kids
        self.nlevel = max(levels)
        ngrid = sum(len(v) for v in levels.values())
        self.ngridmax = ngrid + extra_grids
        self.ncell = NCOARSE + 8 * self.ngridmax
        ngm = self.ngridmax
        son = np.zeros(self.ncell, dtype=np.int32)
        father = np.zeros(ngm, dtype=np.int32)
        nbor = np.zeros(6 * ngm, dtype=np.int32)
        grid_at = {}
        active = {}
        ig = 0
        for l in range(1, self.nlevel + 1):
            ids = []
            for g in levels[l]:
                ig += 1
                grid_at[(l,) + g] = ig
                ids.append(ig)
            active[l] = np.array(ids, dtype=np.int32)
        self.grid_at = grid_at
        self.active = active
        self.level_of_grid = np.zeros(ngm + 1, dtype=np.int32)
        self.grid_pos = np.zeros((ngm + 1, 3), dtype=np.int64)
        for (l, gx, gy, gz), i in grid_at.items():
            self.level_of_grid[i] = l
            self.grid_pos[i] = (gx, gy, gz)

        def cell_index(l, cx, cy, cz):
            n = 2 ** l
            cx, cy, cz = cx % n, cy % n, cz % n
            g = grid_at[(l, cx // 2, cy // 2, cz // 2)]
            ind = 1 + (cx & 1) + 2 * (cy & 1) + 4 * (cz & 1)
            return NCOARSE + (ind - 1) * ngm + g

        self.cell_index = cell_index
        # level 1: the single coarse cell is its father and its own periodic neighbour
        son[NCOARSE - 1] = 1
        father[0] = NCOARSE
        for j in range(6):
            nbor[j * ngm + 0] = NCOARSE
        dirs = [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]
        for l in range(2, self.nlevel + 1):
            for (cx, cy, cz) in levels[l]:
                # grid at level l sits in the father cell (cx, cy, cz) of level l-1
                g = grid_at[(l, cx, cy, cz)]
                fc = cell_index(l - 1, cx, cy, cz)
                father[g - 1] = fc
                son[fc - 1] = g
                for j, (dx, dy, dz) in enumerate(dirs):
                    nbor[j * ngm + g - 1] = cell_index(l - 1, cx + dx, cy + dy, cz + dz)
        self.son, self.father, self.nbor = son, father, nbor

    # -- helpers -----------------------------------------------------------
    def cells_of_level(self, l):
        """(icell array, centre positions (N,3)) of every cell of level l,
        grid-major in active order, ind-minor."""
        ids = self.active[l]
        ngm = self.ngridmax
        out_i, out_x = [], []
        n = 2 ** l
        for ind in range(8):
            ix, iy, iz = ind & 1, (ind >> 1) & 1, ind >> 2
            gp = self.grid_pos[ids]
            c = np.stack([2 * gp[:, 0] + ix, 2 * gp[:, 1] + iy, 2 * gp[:, 2] + iz], axis=1)
            out_i.append(NCOARSE + ind * ngm + ids)
            out_x.append((c + 0.5) / n)
        return np.concatenate(out_i), np.concatenate(out_x)

    def fill(self, fn):
        """Conserved state u[var][icell-1] with fn(positions (N,3), level) ->
        (5, N) evaluated at the centre of every cell of every level."""
        u = np.zeros((5, self.ncell))
        for l in range(1, self.nlevel + 1):
            ic, x = self.cells_of_level(l)
            u[:, ic - 1] = fn(x, l)
        u[:, NCOARSE - 1] = u[:, self.cells_of_level(1)[0] - 1].mean(axis=1)
        return u

    def n_octs(self, l):
        return len(self.active.get(l, []))

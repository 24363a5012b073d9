from specula_amd.base import BaseDataObj
from specula_amd.lib.grids import make_xy


class Lenslet(BaseDataObj):
    """specula/data_objects/lenslet.py:6-39: lenses[i][j] = [x(j), y(i), 2/n]."""

    def __init__(self, n_lenses: int = 1, target_device_idx: int = None, precision: int = None):
        super().__init__(target_device_idx=target_device_idx, precision=precision)
        self.n_lenses = n_lenses
        if n_lenses > 1:
            x, y = make_xy(n_lenses, 1.0)
        else:
            x = [[0.0]]
            y = [[0.0]]
        size = 2.0 / n_lenses
        self._lenses = [[[float(x[i][j]), float(y[i][j]), size] for j in range(n_lenses)]
                        for i in range(n_lenses)]

    @property
    def dimx(self):
        return len(self._lenses)

    @property
    def dimy(self):
        return len(self._lenses[0]) if self._lenses else 0

    def get(self, x, y):
        return self._lenses[x][y]

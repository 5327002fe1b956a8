"""Beam source geometry (host side).

Follows conehead/source.py:5-111: source at (0, -SAD, 0) for gantry 0,
IEC 61217 gantry rotation about z, collimator rotation about the beam axis,
and `transform` = inverse of the basis matrix [v_x v_y v_z] (world -> source).
"""
import numpy as np


class Source:
    """Point source with gantry and collimator rotations.

    Only the "varian_clinac_6MV" model exists (conehead/source.py:9-14); its
    spectrum-weight model is not part of the dose path (the TERMA energy
    weights come from settings['energy_weights']), so `weights` is None.
    """

    def __init__(self, model, sad=np.float32(100)):
        if model != "varian_clinac_6MV":
            raise NotImplementedError("The requested model is not yet"
                                      " implemented.")
        self.weights = None
        self._sad = sad
        self._gantry = 0
        self._collimator = 0
        self._position = np.array([0, -self._sad, 0], dtype=np.float32)
        self.v_x = np.array([1, 0, 0], dtype=np.float32)
        self.v_y = np.array([0, 1, 0], dtype=np.float32)
        self.v_z = np.array([0, 0, 1], dtype=np.float32)
        self.transform = np.linalg.inv(
            np.array([self.v_x, self.v_y, self.v_z]).transpose())

    @property
    def position(self):
        return self._position

    @position.setter
    def position(self, value):
        self._position = value

    @property
    def sad(self):
        return self._sad

    @property
    def gantry(self):
        return self._gantry

    @gantry.setter
    def gantry(self, theta):
        assert 0 <= theta < 360, "Invalid gantry angle"
        self._gantry = theta
        self._update_geometry()

    @property
    def collimator(self):
        return self._collimator

    @collimator.setter
    def collimator(self, theta):
        assert 0 <= theta < 360, "Invalid collimator angle"
        self._collimator = theta
        self._update_geometry()

    def _update_geometry(self):
        # position on the gantry circle (conehead/source.py:79-85); angles
        # are promoted to float64 (NumPy 2 would keep a float32 angle in
        # float32 and put the gantry-0 source 4e-6 cm off axis)
        phi = (90 - float(self._gantry)) % 360
        self.position = np.array([
            float(self.sad) * np.cos(phi * np.pi / 180),
            float(self.sad) * -np.sin(phi * np.pi / 180),
            self.position[2],
        ])
        ex = np.array([1, 0, 0], dtype=np.float32)
        ey = np.array([0, 1, 0], dtype=np.float32)
        ez = np.array([0, 0, 1], dtype=np.float32)
        # collimator about the beam axis, then gantry about z (:92-107)
        c = -float(self._collimator) * np.pi / 180
        rot_c = np.array([[np.cos(c), 0, np.sin(c)],
                          [0, 1, 0],
                          [-np.sin(c), 0, np.cos(c)]])
        g = float(self._gantry) * np.pi / 180
        rot_g = np.array([[np.cos(g), -np.sin(g), 0],
                          [np.sin(g), np.cos(g), 0],
                          [0, 0, 1]])
        self.v_x = np.matmul(rot_g, np.matmul(rot_c, ex))
        self.v_y = np.matmul(rot_g, ey)
        self.v_z = np.matmul(rot_g, np.matmul(rot_c, ez))
        self.transform = np.linalg.inv(
            np.array([self.v_x, self.v_y, self.v_z]).transpose())

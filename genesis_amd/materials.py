"""``gs.materials.Rigid`` (genesis/engine/materials/rigid.py): density and friction of rigid entities."""


class Rigid:
    def __init__(self, rho: float | None = None, friction: float | None = None, needs_coup: bool = True,
                 coup_friction: float = 0.1, coup_softness: float = 0.002, coup_restitution: float = 0.0,
                 gravity_compensation: float = 0.0, **_ignored):
        if friction is not None and not (0.01 <= friction <= 5.0):
            raise ValueError("friction must be in [0.01, 5.0]")
        self.rho = rho
        self.friction = friction
        self.gravity_compensation = gravity_compensation

"""Physics-process plugin base, same contract as the reference's ``PhysProc``
(ocelot/cpbd/physics_proc.py:14-70): the Navigator assigns ``indx0``,
``indx1``, ``s_start``, ``s_stop``, ``start_elem``, ``end_elem``; ``track``
sets ``z0`` and calls ``apply(p_array, dz)`` every ``step`` unit steps;
``prepare(lat)`` runs when the process is added; ``finalize`` at the end."""


class PhysProc:
    def __init__(self, step=1):
        self.step = step
        self.energy = None
        self.indx0 = None
        self.indx1 = None
        self.s_start = None
        self.s_stop = None
        self.start_elem = None
        self.end_elem = None
        self.z0 = None

    def check_step(self):
        # physics_proc.py:41-43 (the reference's condition, kept as is)
        if not isinstance(self.step, (int, float)) and float(self.step).is_integer():
            raise ValueError(f'step must be an integer number, instead {self.step}')

    def prepare(self, lat):
        self.check_step()

    def apply(self, p_array, dz):
        pass

    def finalize(self, *args, **kwargs):
        pass

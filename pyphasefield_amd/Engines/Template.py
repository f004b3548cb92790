"""The engine plugin contract (Engines/Template.py:15-46 of the reference)."""

from ..simulation import Simulation


class Template(Simulation):
    def __init__(self, **kwargs):
        super().__init__(**kwargs)
        # runs before tdb, thermal, fields and boundary conditions are initialised

    def init_tdb_params(self):
        super().init_tdb_params()

    def init_fields(self):
        # runs after tdb and thermal data, before boundary conditions
        pass

    def initialize_fields_and_imported_data(self):
        super().initialize_fields_and_imported_data()

    def just_before_simulating(self):
        super().just_before_simulating()

    def simulation_loop(self):
        # one time step
        pass

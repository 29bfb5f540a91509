"""Residual tracker (reference: residuals/manager.py:2-16): keeps only the Ru
series, seeded with 1.0; ``plateau_detect`` always returns False."""


class ResidualManager:
    class Series:
        def __init__(self):
            self.values = [1.0]

    def __init__(self):
        self.series = {"Ru": ResidualManager.Series()}

    def __call__(self, *args, **kwargs):
        return None

    def add(self, name, value):
        if name == "Ru":
            self.series["Ru"].values.append(value)

    def plateau_detect(self, name, window=10, threshold=-0.005):
        return False

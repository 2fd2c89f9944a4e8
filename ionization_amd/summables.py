"""
Composable terms: potentials, windows and masks combine with ``+``.

Behaviour follows ``ionization/summables.py:6-78``: a ``Summand`` iterates as
itself, ``a + b`` builds the summand's ``summation_class`` holding both, and a
sum calls every member and combines the results (added for potentials,
multiplied for windows and masks, which override ``combine``).
"""

import functools
import operator


class Summand:
    summation_class = None  # set by subclasses

    def __iter__(self):
        yield self

    def __add__(self, other):
        cls = self.summation_class or Sum
        return cls(*self, *other)

    def __call__(self, *args, **kwargs):
        raise NotImplementedError

    def __repr__(self):
        return self.__class__.__name__


class Sum(Summand):
    container_name = "summands"
    combine = staticmethod(lambda values: functools.reduce(operator.add, values))

    def __init__(self, *summands):
        setattr(self, self.container_name, tuple(summands))

    @property
    def _container(self):
        return getattr(self, self.container_name)

    def __iter__(self):
        yield from self._container

    def __getitem__(self, item):
        return self._container[item]

    def __len__(self):
        return len(self._container)

    def __add__(self, other):
        return self.__class__(*self, *other)

    def __call__(self, *args, **kwargs):
        return self.combine([x(*args, **kwargs) for x in self._container])

    def __repr__(self):
        return f"{self.__class__.__name__}({', '.join(repr(x) for x in self._container)})"

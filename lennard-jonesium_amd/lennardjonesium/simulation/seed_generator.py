"""Random seeds (api/seed_generator.hpp:32-47): fresh seeds from the operating
system's entropy source; the default seed is std::mt19937's, 5489."""

import random


class SeedGenerator:
    _rng = random.SystemRandom()

    @staticmethod
    def default_seed() -> int:
        return 5489

    def get(self) -> int:
        return self._rng.getrandbits(32)

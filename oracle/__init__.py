"""CPU oracle of the splat pipeline -- TEST INFRASTRUCTURE ONLY (see mirror.cpp)."""

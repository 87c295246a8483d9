// DVRange is declared with the other vectors (reference: fake_vectors/DVRange.h).
#include "../DVVector.h"

// ORACLE — test infrastructure only (see oracle_skin.cpp header).
#pragma once
#include <cstddef>
#include <cstdint>

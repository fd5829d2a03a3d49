// novaplus -- CLI driver (see driver.hpp).
#include "driver.hpp"

int main(int argc, char **argv) { return lpa_cli::driver_main(argc, argv, "novaplus", 0); }

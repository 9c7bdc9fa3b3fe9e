// jacobi2D: 2D 5-point Jacobi driver with the reference's CLI.
#include "jacobi_cli.hpp"

int main(int argc, char *argv[]) { return cli::run_jacobi(argc, argv, 2); }

// solver_operator_mms: the reference's tools/operator_mms_main.cpp on the GPU.
#include "solver/operator_verification.hpp"

#include <iostream>
#include <stdexcept>
#include <string>

int main(int argc, char** argv) {
  try {
    const int n = argc > 1 ? std::stoi(argv[1]) : 32;
    const solver::OperatorManufacturedSolutionResult r = solver::run_operator_manufactured_solution_case(n);
    std::cout << "resolution: " << r.resolution << '\n';
    std::cout << "domain_length: " << r.domain_length << '\n';
    std::cout << "gradient_l2_error: " << r.gradient_error << '\n';
    std::cout << "divergence_l2_error: " << r.divergence_error << '\n';
    std::cout << "laplacian_l2_error: " << r.laplacian_error << '\n';
    return 0;
  } catch(const std::exception& e) {
    std::cerr << "solver_operator_mms failed: " << e.what() << '\n';
    return 1;
  }
}

// hbtx_hdf5.hpp — the per-particle energy datasets of a subhalo snapshot file,
// as SubhaloSnapshot_t::Save writes them (src/io/subhalo_io.cpp:425-451):
//
//   BindingEnergies    if SaveBoundParticleBindingEnergies
//   PotentialEnergies  if SaveBoundParticlePotentialEnergies
//
// Each is a one-dimensional dataset over the file's subhaloes (in the order of
// the subhalo vector) of variable-length native-float arrays.  Entry i holds the
// first Nbound values of subhalo i's ParticleBindingEnergies /
// ParticlePotentialEnergies, in the order of its bound particles (most bound
// first; docs/outputs/format.md:56-63), which hbtx::RefineParticles and
// hbtx::Unbind fill from the GPU's output.  As in the reference, each vector is
// cleared once the dataset is written.
//
// The caller opens and closes the file (the reference's Save writes its other
// datasets into the same file).  This header needs the HDF5 C library; the rest
// of hbtx does not.  C++11.
#pragma once
#include <hdf5.h>

#include <stdexcept>
#include <string>
#include <vector>

namespace hbtx
{
namespace detail
{
struct BindingEnergiesOf
{
  template <class S>
  std::vector<float> &operator()(S &s) const
  {
    return s.ParticleBindingEnergies;
  }
};
struct PotentialEnergiesOf
{
  template <class S>
  std::vector<float> &operator()(S &s) const
  {
    return s.ParticlePotentialEnergies;
  }
};

// writeHDFmatrix (src/hdf_wrapper.cpp:5-26) for a vector of VLEN float arrays.
// A subhalo whose vector holds fewer than Nbound values is an error here (the
// reference would read past its end).
template <class SubhaloT, class Get>
void write_energy_dataset(hid_t file, std::vector<SubhaloT> &subs, const char *name, Get get)
{
  std::vector<hvl_t> vl(subs.size());
  for (size_t i = 0; i < subs.size(); i++)
  {
    std::vector<float> &v = get(subs[i]);
    const long long nb = (long long)subs[i].Nbound;
    if (nb < 0 || (size_t)nb > v.size())
      throw std::runtime_error(std::string("hbtx: ") + name + ": subhalo " + std::to_string(i) + " has Nbound " +
                               std::to_string(nb) + " but " + std::to_string(v.size()) + " energies");
    vl[i].len = (size_t)nb;
    vl[i].p = v.data();
  }
  const hid_t ftype = H5Tvlen_create(H5T_NATIVE_FLOAT);
  const hsize_t dims[1] = {(hsize_t)subs.size()};
  const hid_t space = H5Screate_simple(1, dims, NULL);
  const hid_t dset = H5Dcreate2(file, name, ftype, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  herr_t status = 0;
  if (dset >= 0 && !subs.empty())
    status = H5Dwrite(dset, ftype, H5S_ALL, H5S_ALL, H5P_DEFAULT, vl.data());
  if (dset >= 0)
    H5Dclose(dset);
  H5Sclose(space);
  H5Tclose(ftype);
  if (dset < 0 || status < 0)
    throw std::runtime_error(std::string("hbtx: error writing dataset ") + name);
  for (size_t i = 0; i < subs.size(); i++)
    get(subs[i]).clear();
}
} // namespace detail

// The two datasets of src/io/subhalo_io.cpp:425-451 into an open HDF5 file
// (or group): binding and potential select which are written, as the
// reference's SaveBoundParticleBindingEnergies / SaveBoundParticlePotentialEnergies.
template <class SubhaloT>
void SaveParticleEnergies(hid_t file, std::vector<SubhaloT> &Subhalos, bool binding, bool potential)
{
  if (binding)
    detail::write_energy_dataset(file, Subhalos, "BindingEnergies", detail::BindingEnergiesOf());
  if (potential)
    detail::write_energy_dataset(file, Subhalos, "PotentialEnergies", detail::PotentialEnergiesOf());
}

} // namespace hbtx

// Image being optimised plus its per-image line-search functions: the
// reference's include/classes/image.cuh:4-31.  newP / evaluateXt point at
// the positivity (particular*) or plain (default*) variants
// (src/functions.cu:4566-4592, src/mfs.cu:794-811).
#pragma once

typedef struct functionMap {
  void (*newP)(float*, float*, double, int);
  void (*evaluateXt)(float*, float*, float*, double, int);
} imageMap;

class Image {
 public:
  Image(float* image, int image_count) : image(image), image_count(image_count) {}
  int getImageCount() { return image_count; }
  float* getImage() { return image; }
  float* getErrorImage() { return error_image; }
  imageMap* getFunctionMapping() { return functionMapping; }
  void setImageCount(int i) { this->image_count = i; }
  void setErrorImage(float* f) { this->error_image = f; }
  void setImage(float* i) { this->image = i; }
  void setFunctionMapping(imageMap* f) { this->functionMapping = f; }

 private:
  float* image;
  int image_count;
  float* error_image = nullptr;
  imageMap* functionMapping = nullptr;
};

// src/functions.cu:4566-4592: positivity on image 0 (floor -eta*MINPIX),
// none on the others.
void defaultNewP(float* p, float* xi, double xmin, int image);
void particularNewP(float* p, float* xi, double xmin, int image);
void defaultEvaluateXt(float* xt, float* pcom, float* xicom, double x, int image);
void particularEvaluateXt(float* xt, float* pcom, float* xicom, double x, int image);
// Builds the mapping of src/mfs.cu:794-811 (owned by the caller, delete[]).
imageMap* makeFunctionMapping(int image_count, bool nopositivity);
